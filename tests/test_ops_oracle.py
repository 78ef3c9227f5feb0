"""Pins of the FP64 restatement of the reference's operator library (include/fvmops.cuh:96-219):
closed forms on a uniform hex box (Green-Gauss exactness for linear fields, closure for
constants) and hand-computed values of the quirks the reference writes."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def box(gr, orc):
    n = 4
    return n, orc.OracleMesh(gr.ElementArrays.of(gr.Elements.box(n, n, n, 2.0, 1.0, 1.0)))


def centres(om):
    return om.cell_centers[:om.n_cells + om.n_ghosts]


def test_constant_fields_give_zero(orc, box):
    _, om = box
    n = om.n_cells + om.n_ghosts
    s, v = np.full(n, 1.7), np.tile([0.3, -1.1, 2.0], (n, 1))
    for op, val in ((orc.OP_SCAL_GRAD_C, s), (orc.OP_BOUND_SCAL_GRAD_C, s), (orc.OP_VEC_GRAD_C, v),
                    (orc.OP_BOUND_VEC_GRAD_C, v), (orc.OP_DIV_C, v)):
        assert np.abs(orc.fvm_op(om, op, val)).max() < 1e-12, op
    # a constant gradient has no Laplacian (closure of every cell)
    g = np.tile([0.5, -2.0, 1.0], (n, 1))
    assert np.abs(orc.fvm_op(om, orc.OP_LAPLACIAN_C, s, grad=g)).max() < 1e-12
    # a constant field carried by a constant velocity: no net convection
    assert np.abs(orc.fvm_op(om, orc.OP_CONV_CONST_SCAL_U, s, phi_const=[0.4, 0.1, -0.2])).max() < 1e-12


def test_linear_fields_are_differentiated_exactly(orc, box):
    """Uniform hexes, ghosts mirrored across the boundary: face interpolation is exact for
    linear fields, so the Green-Gauss operators return the exact gradient and divergence."""
    _, om = box
    x = centres(om)
    a = np.array([0.7, -1.3, 2.2])
    g = orc.fvm_op(om, orc.OP_SCAL_GRAD_C, x @ a)
    assert np.allclose(g, a, rtol=0, atol=1e-12)
    A = np.array([[1.0, 2.0, 0.0], [0.5, -1.0, 3.0], [0.0, 0.25, 2.0]])
    v = x @ A.T                                         # v_r = sum_k A[r, k] x_k
    T = orc.fvm_op(om, orc.OP_VEC_GRAD_C, v).reshape(-1, 3, 3)   # column c = grad(v_c)
    assert np.allclose(T, np.broadcast_to(A, T.shape), rtol=0, atol=1e-12)
    d = orc.fvm_op(om, orc.OP_DIV_C, v)
    assert np.allclose(d, np.trace(A), rtol=0, atol=1e-12)
    # the boundary form sums n (u_j - u_i) ds: for a linear u that is grad u . (sum n d ds)
    gb = orc.fvm_op(om, orc.OP_BOUND_SCAL_GRAD_C, x @ a)
    assert np.isfinite(gb).all()


def test_laplacian_of_a_quadratic(orc, box):
    """u = x^2 + 3 y^2 - z^2 has gradient (2x, 6y, -2z), a linear field, and Laplacian 6."""
    _, om = box
    x = centres(om)
    grad = np.stack([2 * x[:, 0], 6 * x[:, 1], -2 * x[:, 2]], axis=1)
    L = orc.fvm_op(om, orc.OP_LAPLACIAN_C, np.zeros(len(x)), grad=grad)
    assert np.allclose(L, 6.0, rtol=0, atol=1e-11)


def test_convection_hand_values(orc, gr):
    """Two unit cubes side by side (ghosts on every other face, all values zero there): the face
    between them has n = +x from cell 0 and ds = 1; the upwinding term of fvmConvConstScalU uses
    max(|a0|, |a1|) while fvmConvVecU uses max(a0, a1)."""
    om = orc.OracleMesh(gr.ElementArrays.of(gr.Elements.box(2, 1, 1, 2.0, 1.0, 1.0)))
    n = om.n_cells + om.n_ghosts
    s = np.zeros(n)
    s[0], s[1] = 1.0, 3.0
    out = orc.fvm_op(om, orc.OP_CONV_CONST_SCAL_U, s, phi_const=[-2.0, 0.0, 0.0])
    # cell 0 and its +x neighbour: a = -2: ((-2*1) + (-2*3) - 2*(3-1)) * 0.5 = -6; boundary
    # faces: a = +2 (xmin) on the -x face with u_g = 0: ((2*1 + 0) - 2*(0-1)) * 0.5 = 2; the
    # y/z faces have a = 0 and contribute (0 - 0*(0-1)) * 0.5 = 0
    assert out[0] == pytest.approx(-6.0 + 2.0, abs=1e-14)
    phi = np.zeros((n, 3))
    phi[0] = [-1.0, 0, 0]
    phi[1] = [-3.0, 0, 0]
    v = np.zeros((n, 3))
    v[0] = [1.0, 0, 0]
    v[1] = [3.0, 0, 0]
    out = orc.fvm_op(om, orc.OP_CONV_VEC_U, v, phi_field=phi)
    # interior face of cell 0: a0 = -1, a1 = -3, max(a0, a1) = -1:
    # ((1*-1 + 3*-3) - (3-1)*(-1)) * 0.5 = (-10 + 2) * 0.5 = -4; xmin face: a0 = +1, a1 = 0
    # (ghost phi 0), max = 1: ((1*1 + 0) - (0-1)*1) * 0.5 = 1
    assert out[0, 0] == pytest.approx(-4.0 + 1.0, abs=1e-14)


def test_scalar_gradient_weights(orc, gr):
    """fvmScalGradC weights the neighbour by |d0|/|d0 - d1|; on a stretched pair of cells
    (lengths 1 and 3) the face value of (u0, u1) = (0, 4) is 4 * 0.5/2 = 1 from cell 0."""
    e = gr.Elements.create(np.array([[0, 0, 0], [1, 0, 0], [4, 0, 0], [0, 1, 0], [1, 1, 0], [4, 1, 0],
                                     [0, 0, 1], [1, 0, 1], [4, 0, 1], [0, 1, 1], [1, 1, 1], [4, 1, 1]], float),
                           [gr.GR_HEX, gr.GR_HEX], [0, 8, 16],
                           [0, 1, 4, 3, 6, 7, 10, 9, 1, 2, 5, 4, 7, 8, 11, 10], [])
    x, k, s, nd, _ = e.arrays()
    # boundary quads: every face of the two hexes except the shared x = 1 face
    quads = [[0, 3, 9, 6], [2, 5, 11, 8], [0, 1, 7, 6], [1, 2, 8, 7], [3, 4, 10, 9], [4, 5, 11, 10],
             [0, 1, 4, 3], [1, 2, 5, 4], [6, 7, 10, 9], [7, 8, 11, 10]]
    mk = [("b", np.full(len(quads), gr.GR_QUAD, np.uint32), np.arange(len(quads) + 1, dtype=np.uint32) * 4,
           np.array(quads, np.uint32).reshape(-1))]
    om = orc.OracleMesh(gr.ElementArrays.of(gr.Elements.create(x, k, s, nd, mk)))
    u = np.zeros(om.n_cells + om.n_ghosts)
    u[1] = 4.0
    g = orc.fvm_op(om, orc.OP_SCAL_GRAD_C, u)
    # cell 0 (volume 1): the shared face (area 1, n = +x) carries 1; the others carry 0
    assert g[0] == pytest.approx([1.0, 0.0, 0.0], abs=1e-14)
