"""graphite_amd: MI355X (gfx950) implementation of GRAPHITE's second-order unstructured
finite-volume residual path.

The compute and the host runtime live in libgraphite_hip.so (C ABI, include/graphite_hip.h);
this module is a thin Python binding over that ABI for tests, benchmarks and scripting.
Host arrays are numpy; device memory is owned by the library.  Layouts follow the
reference's (Alexsimulation/graphite): scalar fields (n,), vector fields (n, 3), scalar
gradients (n, 3), vector gradients (n, 9) as tensor columns.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._capi import (GR_OP_BOUND_SCAL_GRAD_C, GR_OP_BOUND_VEC_GRAD_C, GR_OP_CONV_CONST_SCAL_U,  # noqa: F401
                    GR_OP_CONV_CONST_SCAL_U2, GR_OP_CONV_SCAL_U, GR_OP_CONV_VEC_U, GR_OP_DIV_C,
                    GR_OP_LAPLACIAN_C, GR_OP_SCAL_GRAD_C, GR_OP_SCAL_GRAD_C_CORRECT, GR_OP_VEC_GRAD_C)
from ._capi import (GR_BC_FARFIELD, GR_BC_SLIP, GR_F32, GR_F64, GR_HEX, GR_LIM_BARTH_JESPERSEN,
                    GR_LIM_MICHALAK, GR_LIM_VENKATAKRISHNAN, GR_OK, GR_PYRAMID, GR_QUAD, GR_TET,
                    GR_TRI, GR_WEDGE, ElementsInfo, MeshSizes, SolveResult, SolverConfig, WarmNorm, WingParams, NsiConfig, load)
from ._capi import (GR_ERR_DEVICE, GR_ERR_DIVERGED, GR_ERR_INVALID, GR_ERR_IO, GR_ERR_NOT_FOUND,  # noqa: F401
                    GR_ERR_TOPOLOGY)

__all__ = ["GraphiteError", "Elements", "Mesh", "Field", "Solver", "device_count", "load",
           "euler_init", "euler_bc_slip", "euler_bc_farfield", "euler_step_prepare", "euler_step",
           "euler_residual"]


class GraphiteError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _check(rc: int) -> None:
    if rc != GR_OK:
        raise GraphiteError(rc, load().gr_last_error().decode())


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def device_count() -> int:
    return int(load().gr_device_count())


def stream_copy_gbps(nbytes: int = 1 << 30, reps: int = 20) -> float:
    """Read+write GB/s of a plain device copy kernel over nbytes (bandwidth reference)."""
    out = C.c_double()
    _check(load().gr_stream_copy_probe(nbytes, reps, C.byref(out)))
    return out.value


# ---- element lists ----------------------------------------------------------------------

class Elements:
    """su2-style element lists (nodes, volume elements, boundary markers)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            load().gr_elements_free(h)
            self._h = None

    @staticmethod
    def create(nodes, kinds, starts, elem_nodes, markers=()):
        """markers: iterable of (name, kinds, starts, nodes)."""
        L = load()
        xyz = _f64(nodes).reshape(-1)
        k, s, n = _u32(kinds), _u32(starts), _u32(elem_nodes)
        h = C.c_void_p()
        _check(L.gr_elements_create(len(xyz) // 3, _p(xyz, C.c_double), len(k), _p(k, C.c_uint32),
                                    _p(s, C.c_uint32), _p(n, C.c_uint32), C.byref(h)))
        e = Elements(h)
        for name, mk, ms, mn in markers:
            mk, ms, mn = _u32(mk), _u32(ms), _u32(mn)
            _check(L.gr_elements_add_marker(h, name.encode(), len(mk), _p(mk, C.c_uint32),
                                            _p(ms, C.c_uint32), _p(mn, C.c_uint32)))
        return e

    @staticmethod
    def box(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, tets=False, jitter=0.0, seed=0):
        h = C.c_void_p()
        _check(load().gr_gen_box(nx, ny, nz, lx, ly, lz, int(tets), jitter, seed, C.byref(h)))
        return Elements(h)

    @staticmethod
    def ogrid(n_theta, n_r, n_z, r0, r1, ratio, dz):
        h = C.c_void_p()
        _check(load().gr_gen_ogrid(n_theta, n_r, n_z, r0, r1, ratio, dz, C.byref(h)))
        return Elements(h)

    @staticmethod
    def cgrid_naca(n_i=384, n_j=96, n_wake=64, thickness=0.12, radius=20.0, first_height=1e-3, dz=0.1):
        h = C.c_void_p()
        _check(load().gr_gen_cgrid_naca(n_i, n_j, n_wake, thickness, radius, first_height, dz, C.byref(h)))
        return Elements(h)

    @staticmethod
    def wing(n_around=88, n_prism=20, n_tet=70, n_span=20, n_cap=16, thickness=0.12, sweep_deg=30.0,
             taper=0.56, semispan=1.5, first_height=3e-3, growth=1.2, radius=25.0, scale=1.0):
        """Swept NACA wing, prism layers + tetrahedra, half-sphere far field (1.07 M cells with
        the defaults); markers wall, farfield, symmetry.  Lengths in root chords times scale."""
        prm = WingParams(n_around, n_prism, n_tet, n_span, n_cap, thickness, sweep_deg, taper, semispan,
                         first_height, growth, radius, scale)
        h = C.c_void_p()
        _check(load().gr_gen_wing(C.byref(prm), C.byref(h)))
        return Elements(h)

    @staticmethod
    def read_su2(path: str):
        h = C.c_void_p()
        _check(load().gr_elements_read_su2(path.encode(), C.byref(h)))
        return Elements(h)

    def write_su2(self, path: str) -> None:
        _check(load().gr_elements_write_su2(self._h, path.encode()))

    def info(self) -> ElementsInfo:
        i = ElementsInfo()
        _check(load().gr_elements_info_get(self._h, C.byref(i)))
        return i

    def arrays(self):
        """(nodes (n,3), kinds, starts, elem_nodes, markers[(name, kinds, starts, nodes)])"""
        L = load()
        i = self.info()
        xyz = np.zeros(3 * i.n_nodes)
        k = np.zeros(i.n_elems, np.uint32)
        s = np.zeros(i.n_elems + 1, np.uint32)
        n = np.zeros(i.n_elem_nodes, np.uint32)
        _check(L.gr_elements_get(self._h, _p(xyz, C.c_double), _p(k, C.c_uint32), _p(s, C.c_uint32),
                                 _p(n, C.c_uint32)))
        markers = []
        for m in range(i.n_markers):
            name = C.create_string_buffer(128)
            cnt, nn = C.c_uint32(), C.c_uint32()
            _check(L.gr_elements_get_marker(self._h, m, name, C.byref(cnt), C.byref(nn), None, None, None))
            mk = np.zeros(cnt.value, np.uint32)
            ms = np.zeros(cnt.value + 1, np.uint32)
            mn = np.zeros(nn.value, np.uint32)
            _check(L.gr_elements_get_marker(self._h, m, None, None, None, _p(mk, C.c_uint32),
                                            _p(ms, C.c_uint32), _p(mn, C.c_uint32)))
            markers.append((name.value.decode(), mk, ms, mn))
        return xyz.reshape(-1, 3), k, s, n, markers


@dataclass
class ElementArrays:
    """Plain arrays of an element list (the form the oracle consumes)."""
    nodes: np.ndarray
    elem_types: np.ndarray
    elem_starts: np.ndarray
    elem_nodes: np.ndarray
    markers: list = field(default_factory=list)

    @staticmethod
    def of(e: Elements) -> "ElementArrays":
        x, k, s, n, m = e.arrays()
        return ElementArrays(x, k, s, n, m)


# ---- mesh ---------------------------------------------------------------------------------

class Mesh:
    """cuMesh equivalent: host CSR connectivity + device metrics."""

    def __init__(self, handle):
        self._h = handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            load().gr_mesh_free(h)
            self._h = None

    @staticmethod
    def build(e: Elements) -> "Mesh":
        h = C.c_void_p()
        _check(load().gr_mesh_build(e._h, C.byref(h)))
        return Mesh(h)

    @staticmethod
    def read_fvm(path: str, scalar_bytes: int = 8) -> "Mesh":
        h = C.c_void_p()
        _check(load().gr_mesh_read_fvm(path.encode(), scalar_bytes, C.byref(h)))
        return Mesh(h)

    @staticmethod
    def from_su2(su2_path: str, fvm_path: str, scalar_bytes: int = 8) -> None:
        _check(load().gr_mesh_from_su2(su2_path.encode(), fvm_path.encode(), scalar_bytes))

    @staticmethod
    def make_1d(n: int, x0: float = 0.0, x1: float = 1.0) -> "Mesh":
        h = C.c_void_p()
        _check(load().gr_mesh_make_1d(n, x0, x1, C.byref(h)))
        return Mesh(h)

    def write_fvm(self, path: str, scalar_bytes: int = 8) -> None:
        _check(load().gr_mesh_write_fvm(self._h, path.encode(), scalar_bytes))

    def sizes(self) -> MeshSizes:
        s = MeshSizes()
        _check(load().gr_mesh_sizes_get(self._h, C.byref(s)))
        return s

    @property
    def n_cells(self) -> int:
        return self.sizes().n_cells

    @property
    def n_ghosts(self) -> int:
        return self.sizes().n_ghosts

    def connectivity(self) -> dict:
        s = self.sizes()
        ncg = s.n_cells + s.n_ghosts
        out = dict(
            nodes=np.zeros(3 * s.n_nodes), face_nodes_starts=np.zeros(s.n_faces + 1, np.uint32),
            face_nodes=np.zeros(s.n_face_nodes, np.uint32), cell_faces_starts=np.zeros(ncg + 1, np.uint32),
            cell_faces=np.zeros(s.n_cell_slots, np.uint32), cell_connects=np.zeros(s.n_cell_slots, np.uint32),
            ghost_cell_ids=np.zeros(s.n_ghosts, np.uint32), ghost_cell_starts=np.zeros(s.n_regions + 1, np.uint32))
        a = out
        _check(load().gr_mesh_get_connectivity(
            self._h, _p(a["nodes"], C.c_double), _p(a["face_nodes_starts"], C.c_uint32),
            _p(a["face_nodes"], C.c_uint32), _p(a["cell_faces_starts"], C.c_uint32),
            _p(a["cell_faces"], C.c_uint32), _p(a["cell_connects"], C.c_uint32),
            _p(a["ghost_cell_ids"], C.c_uint32), _p(a["ghost_cell_starts"], C.c_uint32)))
        out["nodes"] = out["nodes"].reshape(-1, 3)
        return out

    def region_id(self, name: str) -> int:
        r = C.c_uint32()
        _check(load().gr_mesh_region_id(self._h, name.encode(), C.byref(r)))
        return r.value

    def region_names(self) -> list:
        out = []
        for r in range(self.sizes().n_regions):
            b = C.create_string_buffer(128)
            _check(load().gr_mesh_region_name(self._h, r, b))
            out.append(b.value.decode())
        return out

    def renumber_rcm(self):
        s = self.sizes()
        c2o = np.zeros(s.n_cells + s.n_ghosts, np.uint32)
        f2o = np.zeros(s.n_faces, np.uint32)
        _check(load().gr_mesh_renumber_rcm(self._h, _p(c2o, C.c_uint32), _p(f2o, C.c_uint32)))
        return c2o, f2o

    def renumber_hilbert(self):
        s = self.sizes()
        c2o = np.zeros(s.n_cells + s.n_ghosts, np.uint32)
        f2o = np.zeros(s.n_faces, np.uint32)
        _check(load().gr_mesh_renumber_hilbert(self._h, _p(c2o, C.c_uint32), _p(f2o, C.c_uint32)))
        return c2o, f2o

    def face_colors(self):
        s = self.sizes()
        col = np.zeros(s.n_faces, np.uint32)
        n = C.c_uint32()
        _check(load().gr_mesh_face_colors(self._h, _p(col, C.c_uint32), C.byref(n)))
        return col, n.value

    def bandwidth(self) -> int:
        b = C.c_uint32()
        _check(load().gr_mesh_bandwidth(self._h, C.byref(b)))
        return b.value

    def compute_metrics(self) -> None:
        _check(load().gr_mesh_compute_metrics(self._h))

    def metrics(self) -> dict:
        s = self.sizes()
        ncg = s.n_cells + s.n_ghosts
        out = dict(face_centers=np.zeros(3 * s.n_faces), face_normals=np.zeros(3 * s.n_faces),
                   face_areas=np.zeros(s.n_faces), cell_centers=np.zeros(3 * ncg),
                   cell_volumes=np.zeros(ncg), cell_matrices=np.zeros(9 * s.n_cells))
        _check(load().gr_mesh_get_metrics(self._h, *[_p(out[k], C.c_double) for k in (
            "face_centers", "face_normals", "face_areas", "cell_centers", "cell_volumes", "cell_matrices")]))
        out["face_centers"] = out["face_centers"].reshape(-1, 3)
        out["face_normals"] = out["face_normals"].reshape(-1, 3)
        out["cell_centers"] = out["cell_centers"].reshape(-1, 3)
        out["cell_matrices"] = out["cell_matrices"].reshape(-1, 9)
        return out

    def set_corrected_reconstruction(self, on: bool) -> None:
        _check(load().gr_euler_set_corrected_reconstruction(self._h, int(on)))


# ---- fields -------------------------------------------------------------------------------

class Field:
    """FvmField equivalent: value/gradient/limiter over cells and ghosts, on the device."""

    def __init__(self, mesh: Mesh, n_comp: int, dtype: int = GR_F64, name: str = "", grads: bool = True):
        h = C.c_void_p()
        _check(load().gr_field_create(mesh._h, n_comp, dtype, name.encode(), int(grads), C.byref(h)))
        self._h, self.mesh, self.n_comp, self.dtype, self.grads = h, mesh, n_comp, dtype, grads
        s = mesh.sizes()
        self.n = s.n_cells + s.n_ghosts

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            load().gr_field_free(h)
            self._h = None

    def upload(self, value) -> None:
        v = _f64(value).reshape(-1)
        assert v.size == self.n * self.n_comp
        _check(load().gr_field_upload(self._h, _p(v, C.c_double)))

    def download(self, grads: bool = False):
        v = np.zeros(self.n * self.n_comp)
        g = np.zeros(self.n * self.n_comp * 3) if grads else None
        l = np.zeros(self.n * self.n_comp) if grads else None
        _check(load().gr_field_download(self._h, _p(v, C.c_double), _p(g, C.c_double) if grads else None,
                                        _p(l, C.c_double) if grads else None))
        shp = (self.n,) if self.n_comp == 1 else (self.n, 3)
        if not grads:
            return v.reshape(shp)
        return v.reshape(shp), g.reshape(self.n, 3 * self.n_comp), l.reshape(shp)

    def init(self, value) -> None:
        v = _f64(np.atleast_1d(value))
        _check(load().gr_field_init(self._h, _p(v, C.c_double)))

    def reduce_sum(self) -> float:
        out = C.c_double()
        _check(load().gr_field_reduce_sum(self._h, C.byref(out)))
        return out.value

    def compute_gradient(self) -> None:
        _check(load().gr_field_compute_gradient(self._h))

    def compute_limiter(self, limiter: int = GR_LIM_MICHALAK) -> None:
        _check(load().gr_field_compute_limiter(self._h, limiter))

    def gradlim_bc(self) -> None:
        _check(load().gr_field_gradlim_bc(self._h))

    def bc_fixed(self, region: int, value) -> None:
        v = _f64(np.atleast_1d(value))
        _check(load().gr_field_bc_fixed(self._h, region, _p(v, C.c_double)))

    def bc_zero_grad(self, region: int) -> None:
        _check(load().gr_field_bc_zero_grad(self._h, region))

    def operator(self, op: int, phi_const=None, phi_field: "Field | None" = None) -> np.ndarray:
        """Finite-volume operator `op` (GR_OP_*) of every interior cell:
        (n_cells,), (n_cells, 3) or (n_cells, 9)."""
        from ._capi import OP_OUTPUTS
        n = self.mesh.n_cells
        # an unknown op still goes to the library, which rejects it
        width = OP_OUTPUTS[op] if 0 <= op < len(OP_OUTPUTS) else 9
        out = np.zeros(n * width)
        pc = None if phi_const is None else _f64(phi_const)
        _check(load().gr_field_operator(self._h, op, None if pc is None else _p(pc, C.c_double),
                                        None if phi_field is None else phi_field._h, _p(out, C.c_double)))
        return out if width == 1 else out.reshape(n, -1)


def euler_init(mesh, rho, rhou, rhoe, mach, p, T):
    m = _f64(mach)
    _check(load().gr_euler_init(mesh._h, rho._h, rhou._h, rhoe._h, _p(m, C.c_double), p, T))


def euler_bc_slip(mesh, rho, rhou, rhoe, region: str):
    _check(load().gr_euler_bc_slip(mesh._h, rho._h, rhou._h, rhoe._h, region.encode()))


def euler_bc_farfield(mesh, rho, rhou, rhoe, mach, p, T, region: str):
    m = _f64(mach)
    _check(load().gr_euler_bc_farfield(mesh._h, rho._h, rhou._h, rhoe._h, _p(m, C.c_double), p, T,
                                       region.encode()))


def euler_step_prepare(mesh, rho, rhou, rhoe, limiter: int = GR_LIM_MICHALAK):
    _check(load().gr_euler_step_prepare(mesh._h, rho._h, rhou._h, rhoe._h, limiter))


def euler_step(mesh, rho, rhou, rhoe, cfl: float, dt: Field | None = None, error: Field | None = None,
               time_scheme: str = "steady"):
    _check(load().gr_euler_step(mesh._h, rho._h, rhou._h, rhoe._h, cfl, time_scheme.encode(),
                                dt._h if dt else None, error._h if error else None))


def euler_residual(mesh, rho, rhou, rhoe) -> np.ndarray:
    out = np.zeros(5 * mesh.n_cells)
    _check(load().gr_euler_residual(mesh._h, rho._h, rhou._h, rhoe._h, _p(out, C.c_double)))
    return out.reshape(-1, 5)


# ---- solver -------------------------------------------------------------------------------

class Solver:
    """Device-resident Euler iteration (PhysicsEuler::step protocol)."""

    def __init__(self, mesh: Mesh, dtype: int = GR_F64, cfl: float = 0.5, limiter: int = GR_LIM_MICHALAK,
                 corrected: bool = False, use_graph: bool = True):
        cfg = SolverConfig(dtype, cfl, limiter, int(corrected), int(use_graph))
        h = C.c_void_p()
        _check(load().gr_solver_create(mesh._h, C.byref(cfg), C.byref(h)))
        self._h, self.mesh, self.dtype = h, mesh, dtype

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            load().gr_solver_free(h)
            self._h = None

    def set_bc(self, region: str, kind: int, mach=(0.0, 0.0, 0.0), p: float = 0.0, T: float = 0.0) -> None:
        m = _f64(mach)
        _check(load().gr_solver_set_bc(self._h, region.encode(), kind, _p(m, C.c_double), p, T))

    def init_freestream(self, mach, p: float, T: float) -> None:
        m = _f64(mach)
        _check(load().gr_solver_init_freestream(self._h, _p(m, C.c_double), p, T))

    def set_state(self, rho, rhou, rhoe) -> None:
        r, u, e = _f64(rho), _f64(rhou).reshape(-1), _f64(rhoe)
        _check(load().gr_solver_set_state(self._h, _p(r, C.c_double), _p(u, C.c_double), _p(e, C.c_double)))

    def get_state(self):
        s = self.mesh.sizes()
        n = s.n_cells + s.n_ghosts
        r, u, e = np.zeros(n), np.zeros(3 * n), np.zeros(n)
        _check(load().gr_solver_get_state(self._h, _p(r, C.c_double), _p(u, C.c_double), _p(e, C.c_double)))
        return r, u.reshape(-1, 3), e

    def iterate(self, max_iter: int, tol: float = 0.0, reference: float = 0.0) -> SolveResult:
        res = SolveResult()
        _check(load().gr_solver_iterate(self._h, max_iter, tol, reference, C.byref(res)))
        return res

    def history(self, n: int) -> np.ndarray:
        h = np.zeros(max(n, 1))
        _check(load().gr_solver_history(self._h, _p(h, C.c_double), n))
        return h[:n]

    def residual_norm(self) -> float:
        out = C.c_double()
        _check(load().gr_solver_residual_norm(self._h, C.byref(out)))
        return out.value

    KERNELS = ("bc", "grad_lim", "flux_update", "residual_reduce")

    def autotune(self, n_iter: int = 20) -> int:
        """Chooses the tile size (cells per block) whose kernels run fastest on this mesh;
        returns it.  The state is not advanced."""
        c = C.c_uint32()
        _check(load().gr_solver_autotune(self._h, n_iter, C.byref(c)))
        return c.value

    def profile(self, n_iter: int) -> dict:
        """Mean launch duration (ms) of each kernel over n_iter back-to-back launches on the
        current state (not advanced), HIP events on the library's stream."""
        ms = np.zeros(4)
        _check(load().gr_solver_profile(self._h, n_iter, _p(ms, C.c_double)))
        return dict(zip(self.KERNELS, ms.tolist()))

    def kernel_bytes(self) -> dict:
        b = np.zeros(4)
        _check(load().gr_solver_kernel_bytes(self._h, _p(b, C.c_double)))
        return dict(zip(self.KERNELS, b.tolist()))

    def layout_bytes(self) -> dict:
        """Bytes of the arrays one launch of each kernel reads or writes in this library's
        layout (per-slot records, padded state records), each element once."""
        b = np.zeros(4)
        _check(load().gr_solver_layout_bytes(self._h, _p(b, C.c_double)))
        return dict(zip(self.KERNELS, b.tolist()))


class Batch:
    """Up to 64 Euler instances (own far-field Mach vectors) on one mesh, advanced together."""

    KERNELS = ("grad_lim", "flux_update", "residual_reduce")

    def __init__(self, mesh: Mesh, n_inst: int, dtype: int = GR_F64, cfl: float = 0.5,
                 limiter: int = GR_LIM_MICHALAK, corrected: bool = False, use_graph: bool = True):
        cfg = SolverConfig(dtype, cfl, limiter, int(corrected), int(use_graph))
        h = C.c_void_p()
        _check(load().gr_batch_create(mesh._h, n_inst, C.byref(cfg), C.byref(h)))
        self._h, self.mesh, self.n_inst, self.dtype = h, mesh, n_inst, dtype

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            load().gr_batch_free(h)
            self._h = None

    def set_bc(self, region: str, kind: int) -> None:
        _check(load().gr_batch_set_bc(self._h, region.encode(), kind))

    def set_freestream(self, mach, p: float, T: float) -> None:
        m = _f64(mach).reshape(-1)
        assert m.size == 3 * self.n_inst
        _check(load().gr_batch_set_freestream(self._h, _p(m, C.c_double), p, T))

    def set_state(self, inst: int, rho, rhou, rhoe) -> None:
        r, u, e = _f64(rho), _f64(rhou).reshape(-1), _f64(rhoe)
        _check(load().gr_batch_set_state(self._h, inst, _p(r, C.c_double), _p(u, C.c_double), _p(e, C.c_double)))

    def get_state(self, inst: int):
        s = self.mesh.sizes()
        n = s.n_cells + s.n_ghosts
        r, u, e = np.zeros(n), np.zeros(3 * n), np.zeros(n)
        _check(load().gr_batch_get_state(self._h, inst, _p(r, C.c_double), _p(u, C.c_double), _p(e, C.c_double)))
        return r, u.reshape(-1, 3), e

    def iterate(self, max_iter: int, tol: float = 0.0, reference=None):
        """Returns (iterations per instance, converged per instance, batch iterations run)."""
        its = (C.c_int * self.n_inst)()
        conv = (C.c_int * self.n_inst)()
        launched = C.c_int()
        ref = None if reference is None else _f64(reference)
        _check(load().gr_batch_iterate(self._h, max_iter, tol, _p(ref, C.c_double) if ref is not None else None,
                                       its, conv, C.byref(launched)))
        return np.array(its[:], dtype=int), np.array(conv[:], dtype=bool), launched.value

    def iterate_masked(self, max_iter: int, tol: float = 0.0, reference=None, active=None):
        """Iterate the active instances only; returns (iterations, converged, diverged, batch
        iterations run) with divergence reported per instance instead of raised."""
        n = self.n_inst
        its, conv, div = (C.c_int * n)(), (C.c_int * n)(), (C.c_int * n)()
        launched = C.c_int()
        ref = None if reference is None else _f64(reference)
        act = None if active is None else (C.c_int * n)(*[int(bool(a)) for a in active])
        _check(load().gr_batch_iterate_masked(self._h, max_iter, tol,
                                              _p(ref, C.c_double) if ref is not None else None, act, its, conv,
                                              div, C.byref(launched)))
        return (np.array(its[:], dtype=int), np.array(conv[:], dtype=bool), np.array(div[:], dtype=bool),
                launched.value)

    def reset_freestream(self, mask) -> None:
        m = (C.c_int * self.n_inst)(*[int(bool(a)) for a in mask])
        _check(load().gr_batch_reset_freestream(self._h, m))

    def history(self, inst: int, n: int) -> np.ndarray:
        h = np.zeros(max(n, 1))
        _check(load().gr_batch_history(self._h, inst, _p(h, C.c_double), n))
        return h[:n]

    def residual_norms(self) -> np.ndarray:
        out = np.zeros(self.n_inst)
        _check(load().gr_batch_residual_norms(self._h, _p(out, C.c_double)))
        return out

    def profile(self, n_iter: int) -> dict:
        ms = np.zeros(3)
        _check(load().gr_batch_profile(self._h, n_iter, _p(ms, C.c_double)))
        return dict(zip(self.KERNELS, ms.tolist()))

    def kernel_bytes(self) -> dict:
        b = np.zeros(3)
        _check(load().gr_batch_kernel_bytes(self._h, _p(b, C.c_double)))
        return dict(zip(self.KERNELS, b.tolist()))

    def set_graph_iters(self, iters: int) -> None:
        _check(load().gr_batch_set_graph_iters(self._h, iters))

    def autotune(self, n_iter: int = 10) -> dict:
        """Chooses the launch schedule (cells per block, flux occupancy cap, flux instance
        groups, graph length) by timed runs; the state is left unchanged.  Returns the choice."""
        sch = (C.c_int * 4)()
        _check(load().gr_batch_autotune(self._h, n_iter, sch))
        return self._schedule_dict(sch)

    @staticmethod
    def _schedule_dict(sch) -> dict:
        return {"cells_per_block": sch[0], "flux_waves_cap": bool(sch[1]), "graph_iters": sch[2],
                "instance_groups": sch[3]}

    @property
    def schedule(self) -> dict:
        sch = (C.c_int * 4)()
        _check(load().gr_batch_get_schedule(self._h, sch))
        return self._schedule_dict(sch)

    @schedule.setter
    def schedule(self, s: dict) -> None:
        sch = (C.c_int * 4)(int(s["cells_per_block"]), int(bool(s["flux_waves_cap"])), int(s["graph_iters"]),
                            int(s.get("instance_groups", 1)))
        _check(load().gr_batch_set_schedule(self._h, sch))

    def set_states(self, rho, rhou, rhoe) -> None:
        """All instances: rho/rhoe (n_inst, N+G), rhou (n_inst, N+G, 3)."""
        r, u, e = _f64(rho).reshape(-1), _f64(rhou).reshape(-1), _f64(rhoe).reshape(-1)
        _check(load().gr_batch_set_states(self._h, _p(r, C.c_double), _p(u, C.c_double), _p(e, C.c_double)))

    def get_states(self):
        s = self.mesh.sizes()
        n = s.n_cells + s.n_ghosts
        r, u, e = np.zeros((self.n_inst, n)), np.zeros((self.n_inst, n, 3)), np.zeros((self.n_inst, n))
        _check(load().gr_batch_get_states(self._h, _p(r, C.c_double), _p(u, C.c_double), _p(e, C.c_double)))
        return r, u, e

    def warm_rows(self, wall_distance, origin):
        """Raw warm-start inputs (n_inst*N, 6) and targets (n_inst*N, 5) of the current states,
        row = cell * n_inst + instance; centre coordinates compressed about `origin`."""
        wd = _f64(wall_distance)
        o = np.ascontiguousarray(origin, np.float32)
        rows = self.mesh.n_cells * self.n_inst
        x, t = np.zeros((rows, 6), np.float32), np.zeros((rows, 5), np.float32)
        _check(load().gr_batch_warm_rows(self._h, _p(wd, C.c_double), _p(o, C.c_float), _p(x, C.c_float),
                                         _p(t, C.c_float)))
        return x, t

    def warm_start(self, net, wall_distance, norm, max_ratio: float = 20.0):
        """Network prediction of every instance, judged by its residual (see gr_batch_warm_start).
        A prediction's residual is spread over the whole field while the freestream's sits at the
        walls, so a useful prediction can start above the freestream residual: on a 128x32 C-grid,
        predictions at 3-15 times the freestream residual still reached 1e-3 in 4-15 times fewer
        iterations; the default keeps predictions up to 20 times it (a diverging instance still
        restarts from freestream).  Returns (freestream residuals, kept flags)."""
        wd = _f64(wall_distance)
        fs = np.zeros(self.n_inst)
        used = (C.c_int * self.n_inst)()
        _check(load().gr_batch_warm_start(self._h, net._h, _p(wd, C.c_double), C.byref(norm), max_ratio,
                                          _p(fs, C.c_double), used))
        return fs, np.array(used[:], dtype=bool)


def _solver_force(self, region: str) -> np.ndarray:
    F = np.zeros(3)
    _check(load().gr_solver_force(self._h, region.encode(), _p(F, C.c_double)))
    return F


def _batch_force(self, region: str) -> np.ndarray:
    F = np.zeros(3 * self.n_inst)
    _check(load().gr_batch_force(self._h, region.encode(), _p(F, C.c_double)))
    return F.reshape(-1, 3)


def _wall_distance(self, regions) -> np.ndarray:
    names = (C.c_char_p * len(regions))(*[r.encode() for r in regions])
    out = np.zeros(self.n_cells)
    _check(load().gr_mesh_wall_distance(self._h, names, len(regions), _p(out, C.c_double)))
    return out


Solver.force = _solver_force
Batch.force = _batch_force
Mesh.wall_distance = _wall_distance


def lift_drag(force, mach, rho_inf: float, p_inf: float, s_ref: float):
    """(C_L, C_D) of a force for a far field with Mach vector `mach` in the x-y plane:
    drag along the free-stream velocity, lift perpendicular to it in the x-y plane."""
    mach = np.asarray(mach, float)
    c_inf = np.sqrt(1.4 * p_inf / rho_inf)
    u = mach * c_inf
    q = 0.5 * rho_inf * float(u @ u) * s_ref
    ed = u / np.linalg.norm(u)
    el = np.array([-ed[1], ed[0], 0.0])
    return float(force @ el) / q, float(force @ ed) / q


# ---- warm-start network -----------------------------------------------------------------------

def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


class Mlp:
    """4-layer MLP n_in -> 256 -> 256 -> 256 -> n_out (tanh hidden layers) on bfloat16 matrix
    cores with binary32 accumulation; binary32 master parameters in the padded layout of
    include/graphite_hip.h (per layer W [in_pad][out_pad] then b [out_pad])."""

    SHAPES = ((64, 256), (256, 256), (256, 256), (256, 32))

    def __init__(self, n_in: int, n_out: int, seed: int = 0):
        h = C.c_void_p()
        _check(load().gr_mlp_create(n_in, n_out, seed, C.byref(h)))
        self._h, self.n_in, self.n_out = h, n_in, n_out
        n = C.c_size_t()
        _check(load().gr_mlp_n_params(h, C.byref(n)))
        self.n_params = n.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            load().gr_mlp_free(h)
            self._h = None

    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        _check(load().gr_mlp_get_params(self._h, _p(p, C.c_float)))
        return p

    def set_params(self, p) -> None:
        p = _f32(p)
        assert p.size == self.n_params
        _check(load().gr_mlp_set_params(self._h, _p(p, C.c_float)))

    def forward(self, x) -> np.ndarray:
        x = _f32(x)
        assert x.ndim == 2 and x.shape[1] == self.n_in
        y = np.zeros((x.shape[0], self.n_out), np.float32)
        _check(load().gr_mlp_forward(self._h, _p(x, C.c_float), x.shape[0], _p(y, C.c_float)))
        return y

    def loss_grad(self, x, t):
        x, t = _f32(x), _f32(t)
        assert x.shape[1] == self.n_in and t.shape == (x.shape[0], self.n_out)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_double()
        _check(load().gr_mlp_loss_grad(self._h, _p(x, C.c_float), _p(t, C.c_float), x.shape[0],
                                       C.byref(loss), _p(g, C.c_float)))
        return loss.value, g

    def set_data(self, x, t) -> None:
        """Keep a training set resident on the device."""
        x, t = _f32(x), _f32(t)
        assert x.shape[1] == self.n_in and t.shape == (x.shape[0], self.n_out)
        _check(load().gr_mlp_set_data(self._h, _p(x, C.c_float), _p(t, C.c_float), x.shape[0]))

    def train(self, batch: int, steps: int, lr: float, lr_final: float, seed: int = 0) -> np.ndarray:
        """Adam steps over the resident set (permuted passes, cosine learning-rate decay);
        returns the loss of every step."""
        losses = np.zeros(steps)
        _check(load().gr_mlp_train(self._h, batch, steps, lr, lr_final, seed, _p(losses, C.c_double)))
        return losses

    def train_step(self, x, t, lr: float) -> float:
        x, t = _f32(x), _f32(t)
        assert x.shape[1] == self.n_in and t.shape == (x.shape[0], self.n_out)
        loss = C.c_double()
        _check(load().gr_mlp_train_step(self._h, _p(x, C.c_float), _p(t, C.c_float), x.shape[0], lr,
                                        C.byref(loss)))
        return loss.value


def warm_norm(x, t, origin=(0.0, 0.0, 0.0), trust=(2.0, 4.0)) -> WarmNorm:
    """Input/output standardisation of a warm-start training set (mean, std per column); origin
    is the point the rows' centre coordinates were compressed about.  trust: wall distances
    over which the prediction is blended out to the freestream (far-wake cells started from a
    prediction made transonic C-grid flows diverge, those started from freestream do not)."""
    x, t = np.asarray(x, np.float64), np.asarray(t, np.float64)
    n = WarmNorm()
    for k in range(3):
        n.origin[k] = origin[k]
    n.trust[0], n.trust[1] = trust
    xm, xs = x.mean(0), x.std(0)
    tm, ts = t.mean(0), t.std(0)
    for c in range(6):
        n.in_mean[c], n.in_std[c] = xm[c], xs[c] if xs[c] > 1e-12 else 1.0
    for c in range(5):
        n.out_mean[c], n.out_std[c] = tm[c], ts[c] if ts[c] > 1e-12 else 1.0
    return n


def normalise(x, t, norm: WarmNorm):
    """(x - in_mean)/in_std and (t - out_mean)/out_std as binary32."""
    xm, xs = np.array(norm.in_mean[:]), np.array(norm.in_std[:])
    tm, ts = np.array(norm.out_mean[:]), np.array(norm.out_std[:])
    return (((x - xm) / xs).astype(np.float32), ((t - tm) / ts).astype(np.float32))


# ---- incompressible pressure-correction solver ----------------------------------------------

class Nsi:
    """Pressure-correction solver for incompressible flow (constant density 1) on a mesh."""

    def __init__(self, mesh: Mesh, nu: float, dt: float, dtype: int = GR_F64, cg_max: int = 200,
                 cg_tol: float = 1e-6):
        cfg = NsiConfig(dtype, nu, dt, cg_max, cg_tol)
        h = C.c_void_p()
        _check(load().gr_nsi_create(mesh._h, C.byref(cfg), C.byref(h)))
        self._h, self.mesh, self.dtype = h, mesh, dtype

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            load().gr_nsi_free(h)
            self._h = None

    def set_wall(self, region: str, u) -> None:
        v = _f64(u)
        _check(load().gr_nsi_set_wall(self._h, region.encode(), _p(v, C.c_double)))

    def set_state(self, u=None, p=None, F=None) -> None:
        args = []
        for a in (u, p, F):
            if a is None:
                args.append(None)
            else:
                a = _f64(a).reshape(-1)
                args.append(a)
        _check(load().gr_nsi_set_state(self._h, *[None if a is None else _p(a, C.c_double) for a in args]))

    def get_state(self):
        s = self.mesh.sizes()
        n = s.n_cells + s.n_ghosts
        u, p, F = np.zeros((n, 3)), np.zeros(n), np.zeros(max(s.n_faces, 1))
        _check(load().gr_nsi_get_state(self._h, _p(u, C.c_double), _p(p, C.c_double), _p(F, C.c_double)))
        return u, p, F[:s.n_faces]

    def step(self, n: int = 1):
        """Returns (continuity residual, CG iterations) of each step."""
        c = np.zeros(n)
        it = (C.c_int * n)()
        _check(load().gr_nsi_step(self._h, n, _p(c, C.c_double), it))
        return c, np.array(it[:], dtype=int)

    def iterate(self, max_steps: int, tol: float):
        steps, ratio = C.c_int(), C.c_double()
        _check(load().gr_nsi_iterate(self._h, max_steps, tol, C.byref(steps), C.byref(ratio)))
        return steps.value, ratio.value

    def history(self, n: int):
        c = np.zeros(max(n, 1))
        it = (C.c_int * max(n, 1))()
        _check(load().gr_nsi_history(self._h, _p(c, C.c_double), it, n))
        return c[:n], np.array(it[:n], dtype=int)

    def profile(self, n_iter: int = 20) -> dict:
        """Mean durations (ms) of the step's parts and of one CG iteration from back-to-back
        launches (the solve is unchanged)."""
        ms = np.zeros(3)
        _check(load().gr_nsi_profile(self._h, n_iter, _p(ms, C.c_double)))
        return dict(zip(("predictor_graph", "cg_iteration", "correction_graph"), ms.tolist()))

    def cg_bytes(self) -> dict:
        """Algorithmic bytes of one CG iteration's kernel (each array element once per launch):
        it reads x, p, r, w, s and the diagonal and writes x, p, r, w, s (the new r, w, s into the
        other slot), and reads the operator's rows (start, one neighbour index and one
        coefficient per interior neighbour); the block partials each block reduces come from L2
        and are not counted."""
        s = self.mesh.sizes()
        N = s.n_cells
        c = self.mesh.connectivity()
        cfs, cc = c["cell_faces_starts"], c["cell_connects"]
        nnz = int((cc[:cfs[N]] < N).sum())
        S = 8 if self.dtype == GR_F64 else 4
        return {"cg_iteration": float(10 * S * N + 8 * N + 4 * (N + 1) + (4 + S) * nnz)}


# ---- case files and VTU output ----------------------------------------------------------------

def write_vtu(mesh: Mesh, path: str, fields: dict) -> None:
    """ASCII VTU of the interior cells (polyhedra) with cell data; fields: name -> array of
    shape (n, ), (n, 3) or (n, 9) with n >= n_cells."""
    names = (C.c_char_p * len(fields))(*[k.encode() for k in fields])
    arrs = [np.ascontiguousarray(v, np.float64) for v in fields.values()]
    nc = np.array([1 if a.ndim == 1 else a.shape[1] for a in arrs], np.uint32)
    vals = (C.POINTER(C.c_double) * len(arrs))(*[_p(a.reshape(-1), C.c_double) for a in arrs])
    _check(load().gr_write_vtu(mesh._h, path.encode(), len(arrs), names, _p(nc, C.c_uint32), vals))


def check_case(path: str) -> None:
    _check(load().gr_case_check(path.encode()))


def run_case(path: str, verbose: bool = False) -> SolveResult:
    r = SolveResult()
    _check(load().gr_case_run(path.encode(), int(verbose), C.byref(r)))
    return r
