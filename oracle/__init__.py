"""FP64 CPU oracle of GRAPHITE's finite-volume residual path (ctypes wrapper).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the cpu_baseline leg
of bench.py, as the checker.  The product package (graphite_amd) never imports it.

The C sources (oracle_mesh.c, oracle_euler.c) restate the reference's host and kernel
routines; this module only marshals numpy arrays.  Field layouts are the reference's AoS
layouts (include/fvmfields.cuh:16-22): scalar fields (n,), vector fields (n, 3), scalar
gradients (n, 3), vector gradients (n, 9) = columns grad(u_x), grad(u_y), grad(u_z).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# GRAPHITE_ORACLE_LIB selects another build of the oracle (its sanitizer build)
_LIB_PATH = os.environ.get("GRAPHITE_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")

LIM_MICHALAK, LIM_BARTH_JESPERSEN, LIM_VENKATAKRISHNAN = 0, 1, 2
BC_SLIP, BC_FARFIELD = 0, 1

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class _OrcMesh(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_uint32), ("n_faces", C.c_uint32), ("n_cells", C.c_uint32),
        ("n_ghosts", C.c_uint32), ("n_regions", C.c_uint32),
        ("nodes", C.POINTER(C.c_double)),
        ("face_nodes_starts", C.POINTER(C.c_uint32)), ("face_nodes", C.POINTER(C.c_uint32)),
        ("cell_faces_starts", C.POINTER(C.c_uint32)), ("cell_faces", C.POINTER(C.c_uint32)),
        ("cell_connects", C.POINTER(C.c_uint32)),
        ("ghost_cell_ids", C.POINTER(C.c_uint32)), ("ghost_cell_starts", C.POINTER(C.c_uint32)),
        ("face_centers", C.POINTER(C.c_double)), ("face_normals", C.POINTER(C.c_double)),
        ("face_areas", C.POINTER(C.c_double)), ("cell_centers", C.POINTER(C.c_double)),
        ("cell_volumes", C.POINTER(C.c_double)), ("cell_matrices", C.POINTER(C.c_double)),
    ]


class _OrcState(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in (
        "rho", "rhou", "rhoe", "grad_rho", "grad_rhou", "grad_rhoe", "lim_rho", "lim_rhou", "lim_rhoe")]


class _OrcBC(C.Structure):
    _fields_ = [("region", C.c_uint32), ("kind", C.c_int), ("mach", C.c_double * 3),
                ("p", C.c_double), ("T", C.c_double)]


class _NsiWall(C.Structure):
    _fields_ = [("region", C.c_uint32), ("u", C.c_double * 3)]


class _NsiParams(C.Structure):
    _fields_ = [("nu", C.c_double), ("dt", C.c_double), ("cg_max", C.c_int), ("cg_tol", C.c_double)]


def build() -> str:
    """Compile liboracle.so with the committed Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_face_hash.restype = C.c_uint32
        L.orc_face_hash.argtypes = [_u32p, C.c_uint32, C.c_uint32]
        L.orc_build_mesh.restype = C.c_int
        L.orc_build_mesh.argtypes = [C.c_uint32, _f64p, C.c_uint32, _u32p, _u32p, _u32p,
                                     C.c_uint32, _u32p, _u32p, _u32p, _u32p,
                                     C.POINTER(C.POINTER(_OrcMesh))]
        L.orc_mesh_free.argtypes = [C.POINTER(_OrcMesh)]
        L.orc_mesh_metrics.argtypes = [C.POINTER(_OrcMesh)]
        L.orc_mesh_check.restype = None
        L.orc_mesh_check.argtypes = [C.POINTER(_OrcMesh), C.POINTER(C.c_uint32)]
        L.orc_limiter_face.restype = C.c_double
        L.orc_limiter_face.argtypes = [C.c_double, _f64p, C.c_double, C.c_double, _f64p, C.c_int]
        mp = C.POINTER(_OrcMesh)
        L.orc_scalar_gradient.argtypes = [mp, _f64p, _f64p]
        L.orc_vector_gradient.argtypes = [mp, _f64p, _f64p]
        L.orc_scalar_limiter.argtypes = [mp, _f64p, _f64p, _f64p, C.c_int]
        L.orc_vector_limiter.argtypes = [mp, _f64p, _f64p, _f64p, C.c_int]
        L.orc_gradlim_bc.argtypes = [mp, C.c_int, _f64p, _f64p]
        L.orc_bc_fixed.argtypes = [mp, C.c_uint32, C.c_int, _f64p, _f64p]
        L.orc_bc_zero_grad.argtypes = [mp, C.c_uint32, C.c_int, _f64p]
        L.orc_reduce_sum.restype = C.c_double
        L.orc_reduce_sum.argtypes = [_f64p, C.c_uint32]
        L.orc_euler_init.argtypes = [mp, _f64p, _f64p, _f64p, _f64p, C.c_double, C.c_double]
        L.orc_euler_slip_bc.argtypes = [mp, C.c_uint32, _f64p, _f64p, _f64p]
        L.orc_euler_farfield_bc.argtypes = [mp, C.c_uint32, _f64p, _f64p, _f64p, _f64p,
                                            C.c_double, C.c_double]
        L.orc_euler_step_prepare.argtypes = [mp, C.POINTER(_OrcState), C.c_int]
        L.orc_euler_step.argtypes = [mp, C.POINTER(_OrcState), C.c_double, C.c_int,
                                     _f64p, _f64p, _f64p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_euler_iterate.restype = C.c_int
        L.orc_euler_iterate.argtypes = [mp, C.POINTER(_OrcState), C.POINTER(_OrcBC), C.c_int,
                                        C.c_double, C.c_int, C.c_int, C.c_int, C.c_double,
                                        C.c_double, _f64p, C.c_int]
        L.orc_fvm_op.restype = C.c_int
        L.orc_fvm_op.argtypes = [mp, C.c_int, _f64p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, _f64p]
        L.orc_nsi_step.restype = C.c_int
        L.orc_nsi_step.argtypes = [mp, _f64p, _f64p, _f64p, C.c_void_p, C.c_int, C.POINTER(_NsiParams),
                                   C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.orc_nsi_poisson.restype = C.c_int
        L.orc_nsi_poisson.argtypes = [mp, _f64p, _f64p, C.c_int, C.c_double]
        L.orc_mlp_n_params.restype = C.c_size_t
        L.orc_mlp_n_params.argtypes = []
        L.orc_bf16_round.restype = C.c_double
        L.orc_bf16_round.argtypes = [C.c_double]
        L.orc_mlp_forward.restype = C.c_int
        L.orc_mlp_forward.argtypes = [_f64p, C.c_int, C.c_int, _f64p, C.c_int, C.c_int, _f64p]
        L.orc_mlp_loss_grad.restype = C.c_int
        L.orc_mlp_loss_grad.argtypes = [_f64p, C.c_int, C.c_int, _f64p, _f64p, C.c_int, C.c_int,
                                        C.POINTER(C.c_double), _f64p]
        _lib = L
    return _lib


def face_hash(key, table_size: int) -> int:
    k = np.ascontiguousarray(key, dtype=np.uint32)
    return int(lib().orc_face_hash(k, len(k), table_size))


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class OracleMesh:
    """Mesh built and measured by the oracle; arrays are numpy copies."""

    def __init__(self, elements, validate: bool = True):
        """elements: object with nodes (n,3) f64, elem_types, elem_starts, elem_nodes (u32) and
        markers = list of (name, types, starts, nodes).  With validate, a mesh the scheme cannot
        run on (see orc_mesh_check) raises ValueError; `invalid` holds the four counts."""
        L = lib()
        nodes = np.ascontiguousarray(elements.nodes, dtype=np.float64).reshape(-1)
        et = np.ascontiguousarray(elements.elem_types, dtype=np.uint32)
        es = np.ascontiguousarray(elements.elem_starts, dtype=np.uint32)
        en = np.ascontiguousarray(elements.elem_nodes, dtype=np.uint32)
        counts = np.array([len(m[1]) for m in elements.markers], dtype=np.uint32)
        mt = np.ascontiguousarray(np.concatenate([np.asarray(m[1], np.uint32) for m in elements.markers])
                                  if elements.markers else np.zeros(1, np.uint32), dtype=np.uint32)
        ms = np.ascontiguousarray(np.concatenate([np.asarray(m[2], np.uint32) for m in elements.markers])
                                  if elements.markers else np.zeros(1, np.uint32), dtype=np.uint32)
        mn = np.ascontiguousarray(np.concatenate([np.asarray(m[3], np.uint32) for m in elements.markers])
                                  if elements.markers else np.zeros(1, np.uint32), dtype=np.uint32)
        if counts.size == 0:
            counts = np.zeros(1, np.uint32)
        out = C.POINTER(_OrcMesh)()
        rc = L.orc_build_mesh(len(nodes) // 3, nodes, len(et), et, es, en, len(elements.markers),
                              counts, mt, ms, mn, C.byref(out))
        if rc != 0:
            raise ValueError("oracle: inconsistent mesh topology")
        self._p = out
        self.region_names = [m[0] for m in elements.markers]
        L.orc_mesh_metrics(out)
        self._copy()
        counts = (C.c_uint32 * 4)()
        L.orc_mesh_check(out, counts)
        self.invalid = tuple(int(c) for c in counts)
        if validate and any(self.invalid):
            raise ValueError("oracle: invalid mesh (non-positive volumes, faces not separating their cells, "
                             f"singular LSQ matrices, degenerate faces) = {self.invalid}")

    def _copy(self):
        s = self._p.contents
        self.n_nodes, self.n_faces, self.n_cells = s.n_nodes, s.n_faces, s.n_cells
        self.n_ghosts, self.n_regions = s.n_ghosts, s.n_regions
        ncg = self.n_cells + self.n_ghosts
        self.nodes = _arr(s.nodes, 3 * s.n_nodes, np.float64).reshape(-1, 3)
        self.face_nodes_starts = _arr(s.face_nodes_starts, s.n_faces + 1, np.uint32)
        self.face_nodes = _arr(s.face_nodes, int(self.face_nodes_starts[-1]), np.uint32)
        self.cell_faces_starts = _arr(s.cell_faces_starts, ncg + 1, np.uint32)
        nslots = int(self.cell_faces_starts[-1])
        self.cell_faces = _arr(s.cell_faces, nslots, np.uint32)
        self.cell_connects = _arr(s.cell_connects, nslots, np.uint32)
        self.ghost_cell_ids = _arr(s.ghost_cell_ids, s.n_ghosts, np.uint32)
        self.ghost_cell_starts = _arr(s.ghost_cell_starts, s.n_regions + 1, np.uint32)
        self.face_centers = _arr(s.face_centers, 3 * s.n_faces, np.float64).reshape(-1, 3)
        self.face_normals = _arr(s.face_normals, 3 * s.n_faces, np.float64).reshape(-1, 3)
        self.face_areas = _arr(s.face_areas, s.n_faces, np.float64)
        self.cell_centers = _arr(s.cell_centers, 3 * ncg, np.float64).reshape(-1, 3)
        self.cell_volumes = _arr(s.cell_volumes, ncg, np.float64)
        self.cell_matrices = _arr(s.cell_matrices, 9 * s.n_cells, np.float64).reshape(-1, 9)

    def region(self, name: str) -> int:
        return self.region_names.index(name)

    def __del__(self):
        p = getattr(self, "_p", None)
        if p:
            lib().orc_mesh_free(p)
            self._p = None


@dataclass
class EulerState:
    """Reference-layout Euler state (interior + ghost cells)."""
    rho: np.ndarray
    rhou: np.ndarray
    rhoe: np.ndarray
    grad_rho: np.ndarray
    grad_rhou: np.ndarray
    grad_rhoe: np.ndarray
    lim_rho: np.ndarray
    lim_rhou: np.ndarray
    lim_rhoe: np.ndarray

    @staticmethod
    def zeros(n: int) -> "EulerState":
        z = lambda *s: np.zeros(s, dtype=np.float64)
        return EulerState(z(n), z(n, 3), z(n), z(n, 3), z(n, 9), z(n, 3), z(n), z(n, 3), z(n))

    def copy(self) -> "EulerState":
        return EulerState(*(getattr(self, f).copy() for f in self.__dataclass_fields__))

    def _c(self) -> _OrcState:
        st = _OrcState()
        for f in self.__dataclass_fields__:
            a = getattr(self, f)
            assert a.flags.c_contiguous and a.dtype == np.float64
            setattr(st, f, a.ctypes.data_as(C.POINTER(C.c_double)))
        return st


def euler_init(m: OracleMesh, s: EulerState, mach, p: float, T: float) -> None:
    lib().orc_euler_init(m._p, s.rho, s.rhou.reshape(-1), s.rhoe, np.asarray(mach, np.float64), p, T)


def slip_bc(m: OracleMesh, region: int, s: EulerState) -> None:
    lib().orc_euler_slip_bc(m._p, region, s.rho, s.rhou.reshape(-1), s.rhoe)


def farfield_bc(m: OracleMesh, region: int, s: EulerState, mach, p: float, T: float) -> None:
    lib().orc_euler_farfield_bc(m._p, region, s.rho, s.rhou.reshape(-1), s.rhoe,
                                np.asarray(mach, np.float64), p, T)


def step_prepare(m: OracleMesh, s: EulerState, limiter: int = LIM_MICHALAK) -> None:
    st = s._c()
    lib().orc_euler_step_prepare(m._p, C.byref(st), limiter)


def step(m: OracleMesh, s: EulerState, cfl: float, corrected: bool = False):
    """Jacobi step; returns (new rho, rhou, rhoe, err, resid (N,5), dt)."""
    n = m.n_cells + m.n_ghosts
    nr, nu, ne = np.zeros(n), np.zeros((n, 3)), np.zeros(n)
    err, res, dt = np.zeros(m.n_cells), np.zeros((m.n_cells, 5)), np.zeros(m.n_cells)
    st = s._c()
    lib().orc_euler_step(m._p, C.byref(st), cfl, int(corrected), nr, nu.reshape(-1), ne,
                         err.ctypes.data, res.ctypes.data, dt.ctypes.data)
    return nr, nu, ne, err, res, dt


def scalar_gradient(m: OracleMesh, value: np.ndarray) -> np.ndarray:
    g = np.zeros((m.n_cells + m.n_ghosts, 3))
    lib().orc_scalar_gradient(m._p, np.ascontiguousarray(value, np.float64), g.reshape(-1))
    return g


def vector_gradient(m: OracleMesh, value: np.ndarray) -> np.ndarray:
    g = np.zeros((m.n_cells + m.n_ghosts, 9))
    lib().orc_vector_gradient(m._p, np.ascontiguousarray(value, np.float64).reshape(-1), g.reshape(-1))
    return g


def scalar_limiter(m: OracleMesh, value, grad, kind: int = LIM_MICHALAK) -> np.ndarray:
    out = np.zeros(m.n_cells + m.n_ghosts)
    lib().orc_scalar_limiter(m._p, np.ascontiguousarray(value, np.float64),
                             np.ascontiguousarray(grad, np.float64).reshape(-1), out, kind)
    return out


def vector_limiter(m: OracleMesh, value, grad, kind: int = LIM_MICHALAK) -> np.ndarray:
    out = np.zeros((m.n_cells + m.n_ghosts, 3))
    lib().orc_vector_limiter(m._p, np.ascontiguousarray(value, np.float64).reshape(-1),
                             np.ascontiguousarray(grad, np.float64).reshape(-1), out.reshape(-1), kind)
    return out


def limiter_face(u, grad, umin, umax, d, kind: int = LIM_MICHALAK) -> float:
    return float(lib().orc_limiter_face(u, np.asarray(grad, np.float64), umin, umax,
                                        np.asarray(d, np.float64), kind))


def reduce_sum(v: np.ndarray) -> float:
    v = np.ascontiguousarray(v, np.float64)
    return float(lib().orc_reduce_sum(v, len(v)))


def iterate(m: OracleMesh, s: EulerState, bcs, cfl: float, n_iter: int, tol: float = 0.0,
            ref_norm: float = 0.0, limiter: int = LIM_MICHALAK, corrected: bool = False,
            threads: int = 1):
    """Run n_iter iterations in place; bcs = list of (region, kind, mach, p, T).
    Returns (iterations run, history array)."""
    arr = (_OrcBC * max(1, len(bcs)))()
    for i, (region, kind, mach, p, T) in enumerate(bcs):
        arr[i].region, arr[i].kind = region, kind
        for k in range(3):
            arr[i].mach[k] = float(mach[k])
        arr[i].p, arr[i].T = float(p), float(T)
    hist = np.zeros(max(1, n_iter))
    st = s._c()
    n = lib().orc_euler_iterate(m._p, C.byref(st), arr, len(bcs), cfl, limiter, int(corrected),
                                n_iter, tol, ref_norm, hist, threads)
    return n, hist[:n]


# ---- warm-start MLP (oracle_mlp.c) ----

MLP_SHAPES = ((64, 256), (256, 256), (256, 256), (256, 32))  # padded (in, out) per layer


def mlp_n_params() -> int:
    return int(lib().orc_mlp_n_params())


def bf16_round(v: float) -> float:
    return float(lib().orc_bf16_round(float(v)))


def mlp_forward(params, x, n_out: int, emulate_bf16: bool = False) -> np.ndarray:
    """y = MLP(x) in binary64 (optionally with the device's bfloat16 operand rounding)."""
    p = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    M, n_in = x.shape
    y = np.zeros((M, n_out))
    if lib().orc_mlp_forward(p, n_in, n_out, x.reshape(-1), M, int(emulate_bf16), y.reshape(-1)):
        raise ValueError("bad MLP dimensions")
    return y


def mlp_loss_grad(params, x, t, emulate_bf16: bool = False):
    """(mean squared error, gradient in the padded parameter layout)."""
    p = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    t = np.ascontiguousarray(t, np.float64)
    M, n_in = x.shape
    g = np.zeros(mlp_n_params())
    loss = C.c_double()
    if lib().orc_mlp_loss_grad(p, n_in, t.shape[1], x.reshape(-1), t.reshape(-1), M, int(emulate_bf16),
                               C.byref(loss), g):
        raise ValueError("bad MLP dimensions")
    return loss.value, g


# ---- incompressible pressure-correction (oracle_nsi.c) ----

class NsiState:
    """u (N+G, 3), p (N+G,), F (n_faces,) face fluxes from the lower- to the higher-numbered cell."""

    def __init__(self, m: OracleMesh):
        n = m.n_cells + m.n_ghosts
        self.u = np.zeros((n, 3))
        self.p = np.zeros(n)
        self.F = np.zeros(m.n_faces)


def nsi_step(m: OracleMesh, st: NsiState, walls, nu: float, dt: float, cg_max: int, cg_tol: float = 0.0):
    """One pressure-correction step in place; walls = [(region, (ux, uy, uz))].
    Returns (continuity residual of the predicted fluxes, CG iterations)."""
    arr = (_NsiWall * max(1, len(walls)))()
    for i, (r, u) in enumerate(walls):
        arr[i].region = r
        for k in range(3):
            arr[i].u[k] = float(u[k])
    prm = _NsiParams(nu, dt, cg_max, cg_tol)
    cont, its = C.c_double(), C.c_int()
    u = np.ascontiguousarray(st.u).reshape(-1)
    lib().orc_nsi_step(m._p, u, st.p, st.F, C.cast(arr, C.c_void_p), len(walls), C.byref(prm), C.byref(cont),
                       C.byref(its))
    st.u = u.reshape(-1, 3)
    return cont.value, its.value


def nsi_poisson(m: OracleMesh, b, cg_max: int, cg_tol: float = 0.0):
    """x with sum_f a_f (x_i - x_j) = b_i (zero-mean b and x); returns (x, iterations)."""
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(m.n_cells)
    its = lib().orc_nsi_poisson(m._p, b, x, cg_max, cg_tol)
    return x, its


# ---- finite-volume operator library (oracle_ops.c) ----

OP_LAPLACIAN_C, OP_SCAL_GRAD_C, OP_SCAL_GRAD_C_CORRECT, OP_VEC_GRAD_C, OP_BOUND_SCAL_GRAD_C, \
    OP_BOUND_VEC_GRAD_C, OP_DIV_C, OP_CONV_CONST_SCAL_U, OP_CONV_CONST_SCAL_U2, OP_CONV_SCAL_U, OP_CONV_VEC_U = range(11)
OP_OUT = {0: 1, 1: 3, 2: 3, 3: 9, 4: 3, 5: 9, 6: 1, 7: 1, 8: 1, 9: 1, 10: 3}


def fvm_op(m: OracleMesh, op: int, val, grad=None, lim=None, phi_const=None, phi_field=None) -> np.ndarray:
    """Operator `op` of every interior cell: (n_cells,) or (n_cells, 3) or (n_cells, 9)."""
    def ptr(a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, np.float64).reshape(-1)
        keep.append(a)
        return a.ctypes.data_as(C.c_void_p)
    keep = []
    v = np.ascontiguousarray(val, np.float64).reshape(-1)
    out = np.zeros(m.n_cells * OP_OUT[op])
    if lib().orc_fvm_op(m._p, op, v, ptr(grad), ptr(lim), ptr(phi_const), ptr(phi_field), out):
        raise ValueError("unknown operator")
    return out.reshape(m.n_cells, -1).squeeze() if OP_OUT[op] > 1 else out
