"""Sweep chunks replayed as HIP graphs (default) and launched one by one (cdx_set_graphs(ctx, 0))
run the same kernels in the same order: results equal bit for bit."""
import numpy as np
import pytest

from cardillocxx_amd import native, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("solver", ["pj", "gs", "pj_chebyshev", "pj_nesterov"])
def test_graph_replay_equals_direct_launches(ctx, solver):
    prob = scenes.random_problem(nb=600, nc=900, seed=21)
    out = []
    try:
        for graphs in (True, False):
            ctx.set_graphs(graphs)
            ctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
            ctx.set_contacts(*prob.contact_arrays())
            p = native.default_params()
            p.tol_abs = p.tol_rel = 1e-9
            if solver == "gs":
                info = ctx.gs_colored_solve(1e-3, 1.0, p)
            else:
                p.accel = {"pj": native.CDX_ACCEL_NONE, "pj_chebyshev": native.CDX_ACCEL_CHEBYSHEV,
                           "pj_nesterov": native.CDX_ACCEL_NESTEROV}[solver]
                info = ctx.pj_solve(1e-3, 1.0, p)
            out.append((info.iters, info.residual, ctx.velocities(), ctx.impulses()))
    finally:
        ctx.set_graphs(True)
    assert out[0][0] == out[1][0] and out[0][0] > 16
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][2], out[1][2]) and np.array_equal(out[0][3], out[1][3])
