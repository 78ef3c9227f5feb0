"""Pins of the FP64 CPU oracle (CPU only).

The reference ships no tests, fixtures or outputs (and cannot be built here), so each oracle
routine is pinned by what the mathematics fixes: closed-form geometry, exactness of the
least-squares gradient for linear fields, hand-derived limiter values at and between the
branch points, characteristic properties of the boundary conditions, and closed forms of the
face flux (consistency with the physical Euler flux, a two-cell first-order case).
"""
import math

import numpy as np
import pytest

from flowcases import P_INF, T_INF


def mesh_of(gr, orc, e):
    return orc.OracleMesh(gr.ElementArrays.of(e))


# ---- metrics (src/mesh.cu:44-219) ----------------------------------------------------------

def test_unit_cube_hex_metrics(gr, orc):
    om = mesh_of(gr, orc, gr.Elements.box(1, 1, 1))
    assert np.allclose(om.face_areas, 1.0, atol=1e-15)
    assert np.allclose(np.abs(om.face_normals).max(axis=1), 1.0, atol=1e-15)
    assert abs(om.cell_volumes[0] - 1.0) < 1e-15
    assert np.allclose(om.cell_centers[0], 0.5, atol=1e-15)
    # ghosts: centre mirrored through the face, volume 2*A*dx = 1 (src/mesh.cu:167-172)
    assert np.allclose(om.cell_volumes[1:], 1.0, atol=1e-15)
    g = om.cell_centers[1:]
    assert np.allclose(np.sort(np.abs(g - 0.5).max(axis=1)), 1.0, atol=1e-15)


def test_regular_tet_metrics(gr, orc):
    v = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float)
    faces = [[0, 2, 1], [0, 1, 3], [1, 2, 3], [0, 3, 2]]
    mk = [("wall", [5] * 4, [0, 3, 6, 9, 12], [n for f in faces for n in f])]
    om = mesh_of(gr, orc, gr.Elements.create(v, [10], [0, 4], [0, 1, 2, 3], mk))
    a = 2 * math.sqrt(2)
    assert abs(om.cell_volumes[0] - a ** 3 / (6 * math.sqrt(2))) < 1e-14
    assert np.allclose(om.cell_centers[0], 0.0, atol=1e-15)
    assert np.allclose(om.face_areas, math.sqrt(3) / 4 * a * a, atol=1e-14)
    # triangle face centre = its vertex mean
    fn = om.face_nodes.reshape(-1, 3)
    assert np.allclose(om.face_centers, v[fn].mean(axis=1), atol=1e-15)


def test_jittered_tets_volumes_and_closure(gr, orc):
    e = gr.Elements.box(6, 5, 4, tets=True, jitter=0.1, seed=3)
    x, k, s, n, _ = e.arrays()
    om = mesh_of(gr, orc, e)
    tets = n.reshape(-1, 4)
    p = x[tets]
    det = np.linalg.det(np.stack([p[:, 1] - p[:, 0], p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]], axis=1))
    vol = np.abs(det) / 6.0
    N = om.n_cells
    assert np.allclose(om.cell_volumes[:N], vol, rtol=1e-12, atol=0)
    assert abs(om.cell_volumes[:N].sum() - 1.0) < 1e-13
    # sum of outward n*A over a closed cell vanishes
    for i in range(0, N, 37):
        acc = np.zeros(3)
        for q in range(om.cell_faces_starts[i], om.cell_faces_starts[i + 1]):
            f = om.cell_faces[q]
            nn = om.face_normals[f]
            d = om.face_centers[f] - om.cell_centers[i]
            acc += nn * np.sign(nn @ d) * om.face_areas[f]
        assert np.abs(acc).max() < 1e-15
    assert np.allclose(np.linalg.norm(om.face_normals, axis=1), 1.0, atol=1e-15)


def test_lsq_matrix_is_inverse_of_normal_matrix(gr, orc):
    om = mesh_of(gr, orc, gr.Elements.box(4, 4, 3, tets=True, jitter=0.1, seed=8))
    for i in range(om.n_cells):
        A = np.zeros((3, 3))
        for q in range(om.cell_faces_starts[i], om.cell_faces_starts[i + 1]):
            d = om.cell_centers[om.cell_connects[q]] - om.cell_centers[i]
            A += np.outer(d, d) / (d @ d)
        M = om.cell_matrices[i].reshape(3, 3).T   # stored as tensor columns
        assert np.allclose(M @ A, np.eye(3), atol=1e-12)
        assert np.array_equal(M, M.T)


# ---- gradient (src/fvmops.cu:51-111) --------------------------------------------------------

def test_gradient_exact_for_linear_fields(gr, orc):
    om = mesh_of(gr, orc, gr.Elements.box(5, 4, 4, tets=True, jitter=0.1, seed=4))
    x = om.cell_centers
    b = np.array([0.3, -1.7, 2.2])
    u = 1.5 + x @ b
    g = orc.scalar_gradient(om, u)[:om.n_cells]
    assert np.abs(g - b).max() < 1e-12
    B = np.array([[1.0, 2.0, -0.5], [0.0, -3.0, 4.0], [7.0, 0.25, 1.0]])   # row c = grad of u_c
    uv = np.array([0.1, 0.2, 0.3]) + x @ B.T
    gv = orc.vector_gradient(om, uv)[:om.n_cells]
    # layout: entries 3c..3c+2 are grad(u_c) (tensor column c)
    assert np.abs(gv - B.reshape(1, 9)).max() < 1e-11


# ---- limiter (src/fvmops.cu:118-169), hand-derived values -----------------------------------

LIM = [(0, "michalak"), (1, "barth_jespersen"), (2, "venkatakrishnan")]


def test_limiter_uniform_field_is_one(orc):
    for kind, _ in LIM:
        assert orc.limiter_face(1.0, [0, 0, 0], 1.0, 1.0, [0.1, 0, 0], kind) == 1.0


@pytest.mark.parametrize("grad,expected", [
    ((5.0, 0, 0), {0: 0.8, 1: 0.8, 2: 0.8 * (4 + 4) / (4 + 2 + 2)}),    # y = 2
    ((10.0, 0, 0), {0: 0.6, 1: 0.8, 2: 0.8 * 3 / 4}),                   # y = 1
    ((-20.0, 0, 0), {0: 0.35, 1: 0.4, 2: 0.8 * 1.25 / 2.75}),           # y = 0.5 via the lower side
])
def test_limiter_sigma_zero_branch(orc, grad, expected):
    # u=0, umax=1, umin=-1 -> dmm2=4 > 2*|d|^3 = 0.002 -> sigma = 0, result = 0.8*phi(y)
    for kind, _ in LIM:
        got = orc.limiter_face(0.0, grad, -1.0, 1.0, [0.1, 0.0, 0.0], kind)
        assert abs(got - expected[kind]) < 1e-15, kind


def test_limiter_sigma_smooth_branch(orc):
    # |d| = 1 -> k3v = 1; umax-umin = sqrt(1.5) -> dmm2 = 1.5, y = 0.5, sigma = 0.5
    h = math.sqrt(1.5) / 2
    got = orc.limiter_face(0.0, [0, 0, 0], -h, h, [1.0, 0.0, 0.0], 0)
    assert abs(got - (0.5 + 0.5 * 0.8)) < 1e-12
    # break points: dmm2 == k3v -> sigma 1; dmm2 == 2 k3v -> sigma 0 (continuous)
    assert orc.limiter_face(0.0, [3.0, 0, 0], -0.5, 0.5, [1.0, 0, 0], 0) == 1.0
    h2 = math.sqrt(2.0) / 2
    got = orc.limiter_face(0.0, [0.5, 0, 0], -h2, h2, [1.0, 0, 0], 0)
    y = h2 / 0.5
    assert abs(got - 0.8 * (y - 0.25 * y * y)) < 1e-12


def test_limiter_small_projection_is_unlimited(orc):
    assert orc.limiter_face(0.0, [1e-13, 0, 0], -1.0, 1.0, [0.5, 0, 0], 0) == pytest.approx(0.8)


# ---- boundary conditions (src/models/euler.cu:412-520) --------------------------------------

def _freestream(mach):
    rho = P_INF / T_INF
    c = math.sqrt(1.4 * P_INF / rho)
    u = np.array(mach) * c
    return rho, rho * u, P_INF / 0.4 + 0.5 * rho * u @ u


@pytest.mark.parametrize("mach", [(0.5, 0.2, 0.0), (2.0, 0.0, -0.3)])
def test_farfield_keeps_freestream(gr, orc, mach):
    om = mesh_of(gr, orc, gr.Elements.box(2, 2, 2))
    n = om.n_cells + om.n_ghosts
    s = orc.EulerState.zeros(n)
    orc.euler_init(om, s, mach, P_INF, T_INF)
    r0, ru0, e0 = _freestream(mach)
    s.rho[om.n_cells:] = 0
    s.rhou[om.n_cells:] = 0
    s.rhoe[om.n_cells:] = 0
    for r in range(om.n_regions):   # inflow and outflow sides, sub- or supersonic
        orc.farfield_bc(om, r, s, mach, P_INF, T_INF)
    assert np.allclose(s.rho, r0, rtol=1e-14) and np.allclose(s.rhoe, e0, rtol=1e-14)
    assert np.allclose(s.rhou, ru0, rtol=1e-14, atol=1e-14)


def test_farfield_subsonic_characteristics(gr, orc):
    om = mesh_of(gr, orc, gr.Elements.box(2, 2, 2))
    N = om.n_cells
    s = orc.EulerState.zeros(N + om.n_ghosts)
    rho_i, u_i, p_i = 1.2, np.array([0.3, 0.0, 0.0]), 0.9
    s.rho[:N] = rho_i
    s.rhou[:N] = rho_i * u_i
    s.rhoe[:N] = p_i / 0.4 + 0.5 * rho_i * u_i @ u_i
    mach = np.array([0.4, 0.1, 0.0])
    xmin, xmax = om.region("xmin"), om.region("xmax")
    orc.farfield_bc(om, xmin, s, mach, P_INF, T_INF)   # u.n < 0: subsonic inflow
    orc.farfield_bc(om, xmax, s, mach, P_INF, T_INF)   # u.n > 0: subsonic outflow
    gi = range(N + om.ghost_cell_starts[xmin], N + om.ghost_cell_starts[xmin + 1])
    for g in gi:   # inflow: pressure from inside, temperature and Mach vector from outside
        p = 0.4 * (s.rhoe[g] - 0.5 * s.rhou[g] @ s.rhou[g] / s.rho[g])
        assert abs(p - p_i) < 1e-14 and abs(p / s.rho[g] - T_INF) < 1e-14
        c = math.sqrt(1.4 * p / s.rho[g])
        assert np.allclose(s.rhou[g] / s.rho[g], mach * c, atol=1e-14)
    go = range(N + om.ghost_cell_starts[xmax], N + om.ghost_cell_starts[xmax + 1])
    for g in go:   # outflow: pressure from outside, temperature and velocity from inside
        p = 0.4 * (s.rhoe[g] - 0.5 * s.rhou[g] @ s.rhou[g] / s.rho[g])
        assert abs(p - P_INF) < 1e-14 and abs(p / s.rho[g] - p_i / rho_i) < 1e-14
        assert np.allclose(s.rhou[g] / s.rho[g], u_i, atol=1e-15)


def test_slip_mirrors_normal_momentum(gr, orc):
    om = mesh_of(gr, orc, gr.Elements.box(2, 2, 2, tets=True, jitter=0.1, seed=1))
    N = om.n_cells
    rng = np.random.default_rng(0)
    s = orc.EulerState.zeros(N + om.n_ghosts)
    s.rho[:N] = 1 + rng.random(N)
    s.rhou[:N] = rng.standard_normal((N, 3))
    s.rhoe[:N] = 3 + rng.random(N)
    for r in range(om.n_regions):
        orc.slip_bc(om, r, s)
    for g in range(om.n_ghosts):
        c = N + g
        q = om.cell_faces_starts[c]
        o, f = om.cell_connects[q], om.cell_faces[q]
        n = om.face_normals[f]
        assert s.rho[c] == s.rho[o] and s.rhoe[c] == s.rhoe[o]
        assert abs(s.rhou[c] @ n + s.rhou[o] @ n) < 1e-14
        t_c = s.rhou[c] - (s.rhou[c] @ n) * n
        t_o = s.rhou[o] - (s.rhou[o] @ n) * n
        assert np.allclose(t_c, t_o, atol=1e-14)


# ---- flux and step (src/models/euler.cu:236-381) --------------------------------------------

def test_two_cell_first_order_closed_form(gr, orc):
    # two unit hexes, all walls slip, fluid at rest, zero gradients: only pressure on the walls
    # and the scalar dissipation across the shared face remain.
    om = mesh_of(gr, orc, gr.Elements.box(2, 1, 1, lx=2.0))
    n = om.n_cells + om.n_ghosts
    s = orc.EulerState.zeros(n)
    rho, p = np.array([1.0, 1.3]), np.array([1.0, 1.6])
    s.rho[:2], s.rhoe[:2] = rho, p / 0.4
    s.lim_rho[:], s.lim_rhou[:], s.lim_rhoe[:] = 1, 1, 1
    for r in range(om.n_regions):
        orc.slip_bc(om, r, s)
    cfl = 0.7
    nr, nu, ne, err, res, dt = orc.step(om, s, cfl)
    c = np.sqrt(1.4 * p / rho)
    eig = max(c)
    left = 0 if om.cell_centers[0, 0] < om.cell_centers[1, 0] else 1
    right = 1 - left
    for i, o, sgn in ((left, right, 1.0), (right, left, -1.0)):
        exp = np.array([-0.5 * eig * (rho[o] - rho[i]), sgn * 0.5 * (p[o] - p[i]), 0.0, 0.0,
                        -0.5 * eig * (p[o] - p[i]) / 0.4])
        assert np.allclose(res[i], exp, atol=1e-15), (i, res[i], exp)
        assert abs(dt[i] - cfl / (5 * c[i] + eig)) < 1e-15
        assert abs(err[i] - res[i] @ res[i]) < 1e-15
        assert abs(nr[i] - (rho[i] - res[i, 0] * dt[i])) < 1e-15


def _linear_state(om, sym):
    x = om.cell_centers
    rho = 1.3 + x @ np.array([0.05, -0.02, 0.03])
    G = np.array([[0.02, 0.01, 0.0], [0.01, -0.03, 0.02], [0.0, 0.02, 0.01]])   # symmetric
    if not sym:
        G = G + np.array([[0, 0.04, 0], [0, 0, 0], [-0.03, 0, 0]])
    rhou = np.array([0.6, 0.1, -0.05]) + x @ G.T
    rhoe = 3.0 + x @ np.array([0.1, 0.02, -0.04])
    return rho, rhou, rhoe


def _textbook_residual(om, rho, rhou, rhoe_lin_coeffs):
    """sum_f F(U(fc)) . n A / V with the linear field evaluated at each face centre."""
    R = np.zeros((om.n_cells, 5))
    for i in range(om.n_cells):
        for q in range(om.cell_faces_starts[i], om.cell_faces_starts[i + 1]):
            f = om.cell_faces[q]
            fc = om.face_centers[f]
            n = om.face_normals[f] * np.sign(om.face_normals[f] @ (fc - om.cell_centers[i]))
            r, ru, e = rhoe_lin_coeffs(fc)
            u = ru / r
            p = 0.4 * (e - 0.5 * ru @ ru / r)
            V = u @ n
            F = np.concatenate([[r * V], ru * V + p * n, [(e + p) * V]])
            R[i] += F * om.face_areas[f] / om.cell_volumes[i]
    return R


@pytest.mark.parametrize("sym", [True, False])
def test_linear_field_flux_matches_physical_flux(gr, orc, sym):
    # Exact gradients and no limiting make both reconstructions meet at each face centre, so
    # the dissipation vanishes and R is the physical flux balance -- except for the reference's
    # transposed momentum product (src/models/euler.cu:252-253), which only agrees with
    # grad(u).d when the momentum gradient is symmetric.
    om = mesh_of(gr, orc, gr.Elements.box(3, 3, 3, tets=True, jitter=0.1, seed=12))
    rho, rhou, rhoe = _linear_state(om, sym)
    s = orc.EulerState.zeros(om.n_cells + om.n_ghosts)
    s.rho[:], s.rhou[:], s.rhoe[:] = rho, rhou, rhoe
    s.grad_rho[:] = orc.scalar_gradient(om, rho)
    s.grad_rhou[:] = orc.vector_gradient(om, rhou)
    s.grad_rhoe[:] = orc.scalar_gradient(om, rhoe)
    s.lim_rho[:], s.lim_rhou[:], s.lim_rhoe[:] = 1, 1, 1
    N = om.n_cells
    for a in (s.grad_rho, s.grad_rhou, s.grad_rhoe):   # ghosts carry their cell's gradient
        a[N:] = a[om.cell_connects[om.cell_faces_starts[N:-1]]]

    def at(pt):
        return (1.3 + pt @ np.array([0.05, -0.02, 0.03]),
                np.array([0.6, 0.1, -0.05]) + pt @ _linear_G(sym).T,
                3.0 + pt @ np.array([0.1, 0.02, -0.04]))

    exp = _textbook_residual(om, rho, rhou, at)
    _, _, _, _, res_corr, _ = orc.step(om, s, 0.5, corrected=True)
    _, _, _, _, res_ref, _ = orc.step(om, s, 0.5, corrected=False)
    scale = np.abs(exp).max() + 1.0
    assert np.abs(res_corr - exp).max() < 1e-11 * scale
    if sym:
        assert np.abs(res_ref - exp).max() < 1e-11 * scale
    else:
        assert np.abs(res_ref - exp).max() > 1e-6


def _linear_G(sym):
    G = np.array([[0.02, 0.01, 0.0], [0.01, -0.03, 0.02], [0.0, 0.02, 0.01]])
    if not sym:
        G = G + np.array([[0, 0.04, 0], [0, 0, 0], [-0.03, 0, 0]])
    return G


def test_freestream_preservation(gr, orc):
    om = mesh_of(gr, orc, gr.Elements.box(8, 8, 8, tets=True, jitter=0.1, seed=2024))
    s = orc.EulerState.zeros(om.n_cells + om.n_ghosts)
    mach = (0.5, 0.0, 0.0)
    orc.euler_init(om, s, mach, P_INF, T_INF)
    for r in range(om.n_regions):
        orc.farfield_bc(om, r, s, mach, P_INF, T_INF)
    orc.step_prepare(om, s)
    _, _, _, _, res, _ = orc.step(om, s, 0.5)
    # |R| <= C eps |F| max(A/V)
    scale = (P_INF / 0.4 + 1) * np.max(om.face_areas) / np.min(om.cell_volumes[:om.n_cells])
    assert np.abs(res).max() < 64 * np.finfo(float).eps * scale


def test_reduce_sum_is_compensated(orc):
    rng = np.random.default_rng(3)
    v = np.concatenate([rng.random(10000) * 1e-8, [1e8], rng.random(10000)])
    assert orc.reduce_sum(v) == math.fsum(v)
