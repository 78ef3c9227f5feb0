"""Transonic swept wing (1.07 M prisms, pyramids and tetrahedra): the mixed-precision solve
reaches a 1e-3 residual drop, and its lift and drag are those of the binary64 solve."""
import math

import numpy as np
import pytest

from flowcases import P_INF, T_INF

pytestmark = pytest.mark.gpu

SCALE = 0.01   # lengths in units of 100 chords: the limiter's threshold |d|^3 is dimensional


@pytest.fixture(scope="module")
def wing(gr):
    m = gr.Mesh.build(gr.Elements.wing(scale=SCALE))
    m.renumber_rcm()
    m.compute_metrics()
    return m


def solve(gr, m, dtype):
    a = math.radians(2.5)
    mach = (0.8 * math.cos(a), 0.8 * math.sin(a), 0.0)
    sv = gr.Solver(m, dtype=dtype, cfl=0.5)
    for name in ("wall", "symmetry"):
        sv.set_bc(name, gr.GR_BC_SLIP, mach, P_INF, T_INF)
    sv.set_bc("farfield", gr.GR_BC_FARFIELD, mach, P_INF, T_INF)
    sv.init_freestream(mach, P_INF, T_INF)
    r = sv.iterate(5000, tol=1e-3)
    s_ref = 1.5 * 0.5 * (1.0 + 0.56) * SCALE * SCALE
    cl, cd = gr.lift_drag(sv.force("wall"), mach, P_INF / T_INF, P_INF, s_ref)
    return r, cl, cd


def test_wing_reaches_1e3_in_both_precisions(gr, wing):
    r32, cl32, cd32 = solve(gr, wing, gr.GR_F32)
    r64, cl64, cd64 = solve(gr, wing, gr.GR_F64)
    print(f"mixed: {r32.iterations} iterations, C_L {cl32:.6f} C_D {cd32:.6f}; "
          f"binary64: {r64.iterations} iterations, C_L {cl64:.6f} C_D {cd64:.6f}")
    assert r32.converged and r64.converged
    assert r32.iterations < 3000 and r64.iterations < 3000
    assert 0.05 < cl64 < 1.0 and 0.0 < cd64 < 0.2
    # measured: 2e-4 (C_L) and 4e-4 (C_D) relative
    assert abs(cl32 - cl64) < 1e-3 * abs(cl64) and abs(cd32 - cd64) < 2e-3 * abs(cd64)
