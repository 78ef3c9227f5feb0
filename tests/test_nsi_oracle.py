"""Pins of the FP64 oracle of the incompressible pressure-correction step (no reference
arithmetic exists for it): closed forms on a uniform hex box, a manufactured Poisson solution,
exactness of the projection, uniform-flow preservation and the cavity's mirror symmetry."""
import numpy as np
import pytest

NU = 1.0 / 400.0


@pytest.fixture(scope="module")
def box(gr, orc):
    n = 6
    return n, orc.OracleMesh(gr.ElementArrays.of(gr.Elements.box(n, n, n)))


def cell(n, i, j, k):
    return i + n * (j + n * k)


def test_first_step_closed_form(orc, box):
    """From rest, with the lid (ymax) moving at U and no pressure correction, the top layer
    gains u_x = 2 dt nu U / h^2 by wall shear and every other value stays zero; the fluxes
    between top-layer cells are h^2 u_x; the continuity residual is 2 n h^2 u_x."""
    n, om = box
    h, dt, U = 1.0 / n, 0.01, 1.0
    st = orc.NsiState(om)
    cont, its = orc.nsi_step(om, st, [(om.region("ymax"), (U, 0, 0))], NU, dt, cg_max=0)
    assert its == 0
    us = 2 * dt * NU * U / h ** 2
    exp = np.zeros((om.n_cells, 3))
    for k in range(n):
        for i in range(n):
            exp[cell(n, i, n - 1, k), 0] = us
    assert np.allclose(st.u[:om.n_cells], exp, rtol=1e-13, atol=1e-16)
    assert not st.p.any()
    assert cont == pytest.approx(2 * n * h * h * us, rel=1e-13)
    # flux through the face between x-neighbours of the top layer, lower to higher cell
    cfs, cf, cc = om.cell_faces_starts, om.cell_faces, om.cell_connects
    c0, c1 = cell(n, 2, n - 1, 3), cell(n, 3, n - 1, 3)
    f = [cf[q] for q in range(cfs[c0], cfs[c0 + 1]) if cc[q] == c1][0]
    assert st.F[f] == pytest.approx(h * h * us, rel=1e-13)
    nz = np.count_nonzero(np.abs(st.F) > 1e-18)
    assert nz == (n - 1) * n                 # only the top layer's x-faces carry flux


def test_poisson_manufactured(orc, box):
    """On a uniform box a_f = A/D = h, so b = h sum_j (x_i - x_j) for a chosen x; CG recovers it."""
    n, om = box
    h = 1.0 / n
    rng = np.random.default_rng(3)
    x = rng.standard_normal((n, n, n))
    x -= x.mean()
    b = np.zeros_like(x)
    for ax in range(3):
        for sgn in (1, -1):
            nb = np.roll(x, sgn, axis=ax)
            mask = np.ones_like(x, bool)
            idx = [slice(None)] * 3
            idx[ax] = 0 if sgn == 1 else n - 1       # rolled-in values are not neighbours
            mask[tuple(idx)] = False
            b += np.where(mask, h * (x - nb), 0.0)
    # cell order i + n (j + n k) is the C order of an array indexed [k, j, i]
    xs, its = orc.nsi_poisson(om, b.reshape(-1), 500, 1e-14)
    assert its < 500
    assert np.allclose(xs, x.reshape(-1), rtol=0, atol=1e-10)


def _box_operator(n):
    """The Poisson operator of the uniform n^3 box as a dense matrix: (A x)_i = h sum_j (x_i -
    x_j) over face neighbours j (a_f = A/D = h^2/h); cell order i + n (j + n k)."""
    h, N = 1.0 / n, n ** 3
    A = np.zeros((N, N))
    for k in range(n):
        for j in range(n):
            for i in range(n):
                c = cell(n, i, j, k)
                for di, dj, dk in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                    ii, jj, kk = i + di, j + dj, k + dk
                    if 0 <= ii < n and 0 <= jj < n and 0 <= kk < n:
                        A[c, c] += h
                        A[c, cell(n, ii, jj, kk)] -= h
    return A


@pytest.mark.parametrize("iters", [1, 2, 3, 7, 12])
def test_poisson_iterates_equal_textbook_pcg(orc, box, iters):
    """The oracle's CG (Chronopoulos-Gear recurrences) has, in exact arithmetic, the iterates of
    the textbook Jacobi-preconditioned CG (one residual-norm product and one search-direction
    product per iteration); after k iterations from x = 0 both agree to rounding.  A wrong beta
    or alpha recurrence still converges eventually but departs here within a few iterations."""
    n, om = box
    A = _box_operator(n)
    rng = np.random.default_rng(11)
    b = rng.standard_normal(n ** 3)
    b -= b.mean()
    dinv = 1.0 / np.diag(A)
    x, r = np.zeros_like(b), b.copy()
    z = dinv * r
    d, rz = z.copy(), r @ z
    for _ in range(iters):
        q = A @ d
        alpha = rz / (d @ q)
        x += alpha * d
        r -= alpha * q
        z = dinv * r
        rz1 = r @ z
        d = z + (rz1 / rz) * d
        rz = rz1
    xs, its = orc.nsi_poisson(om, b, iters, 0.0)
    assert its == iters
    x -= x.mean()
    assert np.abs(xs - x).max() <= 1e-11 * np.abs(x).max()


def test_projection_is_exact(orc, box):
    """After a step solved to a tight tolerance every cell's net outflow vanishes."""
    n, om = box
    st = orc.NsiState(om)
    for _ in range(3):
        orc.nsi_step(om, st, [(om.region("ymax"), (1.0, 0, 0))], NU, 0.02, cg_max=500, cg_tol=1e-14)
    cfs, cf, cc = om.cell_faces_starts, om.cell_faces, om.cell_connects
    worst = 0.0
    for i in range(om.n_cells):
        s = sum(st.F[cf[q]] if i < cc[q] else -st.F[cf[q]] for q in range(cfs[i], cfs[i + 1]))
        worst = max(worst, abs(s))
    assert worst < 1e-12 * np.abs(st.F).max()


def test_uniform_flow_is_preserved(gr, orc):
    """Every wall moving with the flow: a uniform velocity stays uniform and p stays zero."""
    om = orc.OracleMesh(gr.ElementArrays.of(gr.Elements.box(4, 3, 5, tets=True, jitter=0.1, seed=2)))
    U = np.array([0.3, -0.2, 0.1])
    st = orc.NsiState(om)
    st.u[:] = U
    # fluxes of the uniform field, lower- to higher-numbered cell
    cfs, cf, cc = om.cell_faces_starts, om.cell_faces, om.cell_connects
    for i in range(om.n_cells):
        for q in range(cfs[i], cfs[i + 1]):
            j, f = cc[q], cf[q]
            if i < j:
                nrm = om.face_normals[f] * np.sign(om.face_normals[f] @ (om.cell_centers[j] - om.cell_centers[i]))
                st.F[f] = om.face_areas[f] * (U @ nrm)
    walls = [(r, U) for r in range(len(om.region_names))]
    for _ in range(3):
        cont, _ = orc.nsi_step(om, st, walls, NU, 0.05, cg_max=100, cg_tol=1e-13)
    assert np.allclose(st.u[:om.n_cells], U, rtol=0, atol=1e-13)
    assert np.abs(st.p[:om.n_cells]).max() < 1e-12


def test_cavity_mirror_symmetry(orc, box):
    """The lid moves along x, so the flow is symmetric about the mid-plane z = 1/2."""
    n, om = box
    st = orc.NsiState(om)
    for _ in range(8):
        orc.nsi_step(om, st, [(om.region("ymax"), (1.0, 0, 0))], NU, 0.02, cg_max=200, cg_tol=1e-13)
    u = st.u[:om.n_cells].reshape(n, n, n, 3)     # [k, j, i]
    m = u[::-1]
    assert np.abs(u).max() > 1e-3
    assert np.allclose(u[..., 0], m[..., 0], rtol=0, atol=1e-12)
    assert np.allclose(u[..., 1], m[..., 1], rtol=0, atol=1e-12)
    assert np.allclose(u[..., 2], -m[..., 2], rtol=0, atol=1e-12)
