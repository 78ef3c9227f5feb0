"""The operator library on the device against its FP64 oracle (bit for bit in binary64)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[False, True], ids=["hex", "tet"])
def setup(request, gr, orc):
    tets = request.param
    e = gr.Elements.box(5, 4, 3, 1.0, 0.8, 0.6, tets=tets, jitter=0.12 if tets else 0.0, seed=7)
    m = gr.Mesh.build(e)
    m.compute_metrics()
    om = orc.OracleMesh(gr.ElementArrays.of(e))
    n = om.n_cells + om.n_ghosts
    rng = np.random.default_rng(11)
    fields = {}
    for name, nc in (("s", 1), ("v", 3), ("phi", 3)):
        f = gr.Field(m, nc, gr.GR_F64, name)
        f.upload(rng.uniform(0.5, 1.5, (n, nc)).reshape(-1) if nc > 1 else rng.uniform(0.5, 1.5, n))
        f.compute_gradient()
        f.compute_limiter()
        f.gradlim_bc()
        fields[name] = (f, f.download(grads=True))
    return m, om, fields


OPS = ["LAPLACIAN_C", "SCAL_GRAD_C", "SCAL_GRAD_C_CORRECT", "VEC_GRAD_C", "BOUND_SCAL_GRAD_C",
       "BOUND_VEC_GRAD_C", "DIV_C", "CONV_CONST_SCAL_U", "CONV_CONST_SCAL_U2", "CONV_SCAL_U", "CONV_VEC_U"]
VECTOR_IN = {"VEC_GRAD_C", "BOUND_VEC_GRAD_C", "DIV_C", "CONV_VEC_U"}


@pytest.mark.parametrize("name", OPS)
def test_operator_matches_oracle(gr, orc, setup, name):
    m, om, fields = setup
    op = getattr(gr, "GR_OP_" + name)
    f, (val, grad, lim) = fields["v" if name in VECTOR_IN else "s"]
    phi_f, (pval, _, _) = fields["phi"]
    pc = [0.3, -0.7, 0.45]
    dev = f.operator(op, phi_const=pc, phi_field=phi_f)
    ref = orc.fvm_op(om, op, val, grad=grad, lim=lim, phi_const=pc, phi_field=pval)
    assert dev.shape == ref.shape
    assert np.array_equal(dev, ref), np.abs(dev - ref).max()


def test_operator_argument_errors(gr, setup):
    m, om, fields = setup
    s = fields["s"][0]
    with pytest.raises(gr.GraphiteError):
        s.operator(gr.GR_OP_DIV_C)                   # needs a vector field
    with pytest.raises(gr.GraphiteError):
        s.operator(gr.GR_OP_CONV_CONST_SCAL_U)       # needs the constant velocity
    with pytest.raises(gr.GraphiteError):
        s.operator(99)
