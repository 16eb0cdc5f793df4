"""cdx_pj_solve_noalloc (the allocation-free projected Jacobi) against cdx_pj_solve.

The allocation-free entry point runs the same kernels on the same data in the same order, its
incidence lists built on the device in the same (contact) order as the context path builds them on
the host, so velocities, impulses, iteration counts and residuals are equal bit for bit.  Two calls
in a row (the second on a smaller problem, so the shared scratch array holds the first call's
contents) must both match; the error paths report the context path's statuses.
"""
import numpy as np
import pytest
import torch

from cardillocxx_amd import native, scenes

pytestmark = pytest.mark.gpu


def to_dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype).reshape(-1).copy()).to("cuda:0")


def dev_problem(prob, with_fgyro=True, with_warm=True):
    keep = {
        "dof": to_dev(prob.dof, np.int32), "minv": to_dev(prob.minv, np.float64), "v": to_dev(prob.v, np.float64),
        "fext": to_dev(prob.fext, np.float64), "body_a": to_dev(prob.body_a, np.int32),
        "body_b": to_dev(prob.body_b, np.int32), "mu": to_dev(prob.mu, np.float64),
        "frame": to_dev(prob.frame, np.float64), "angA": to_dev(prob.angA, np.float64),
        "angB": to_dev(prob.angB, np.float64), "rest": to_dev(prob.rest, np.float64),
        "vbias": to_dev(prob.vbias, np.float64)}
    if with_fgyro:
        keep["fgyro"] = to_dev(prob.fgyro, np.float64)
    if with_warm:
        keep["warm"] = to_dev(prob.warm, np.float64)
    return keep


def noalloc_solve(prob, params, dt=1e-3, **kw):
    keep = dev_problem(prob, **kw)
    v_out = torch.full((prob.nv,), float("nan"), dtype=torch.float64, device="cuda:0")
    imp_out = torch.full((max(prob.nc, 1) * 3,), float("nan"), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream().cuda_stream
    info = native.pj_solve_noalloc(prob.nb, prob.nc, {k: t.data_ptr() for k, t in keep.items()}, dt, 1.0, params,
                                   v_out.data_ptr(), imp_out.data_ptr(), stream)
    torch.cuda.synchronize()
    return v_out.cpu().numpy(), imp_out.cpu().numpy()[:3 * prob.nc].reshape(-1, 3), info


def context_solve(ctx, prob, params, dt=1e-3, with_fgyro=True, with_warm=True):
    ctx.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro if with_fgyro else None, prob.q)
    warm = prob.warm if with_warm else np.zeros_like(prob.warm)
    ctx.set_contacts(prob.body_a, prob.body_b, prob.mu, prob.frame, prob.angA, prob.angB, prob.rest, prob.vbias, warm)
    info = ctx.pj_solve(dt, 1.0, params)
    return ctx.velocities(), ctx.impulses(), info


def params(**kw):
    p = native.default_params()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


CASES = [
    ("random_1000x1500_tol", lambda: scenes.random_problem(nb=1000, nc=1500, seed=6, rest=0.3),
     dict(tol_abs=1e-8, tol_rel=1e-8)),
    ("random_300x450_fixed", lambda: scenes.random_problem(nb=300, nc=450, seed=5, p_point=0.2),
     dict(max_iter=77, tol_abs=0.0, tol_rel=0.0)),
]


def test_two_calls_match_the_context_solve_bitwise(ctx):
    big = scenes.sphere_pile(n=3000, settle_steps=40, box=(1.0, 1.0, 2.0), seed=2).problem()
    small = scenes.random_problem(nb=120, nc=200, seed=9, p_point=0.2, rest=0.2)
    assert big.nc > 4 * small.nc
    for prob, prm in ((big, params(tol_abs=1e-8, tol_rel=1e-8, max_iter=3000)), (small, params(tol_abs=1e-9, tol_rel=1e-9))):
        v0, i0, inf0 = context_solve(ctx, prob, prm)
        v1, i1, inf1 = noalloc_solve(prob, prm)
        print(f"nb {prob.nb} nc {prob.nc}: {inf1.iters} sweeps, residual {inf1.residual:.3e}")
        assert inf1.iters == inf0.iters and inf1.nrows == inf0.nrows
        assert inf1.residual == inf0.residual
        assert np.array_equal(v1, v0) and np.array_equal(i1, i0)


@pytest.mark.parametrize("name,build,kw", CASES, ids=[c[0] for c in CASES])
def test_matches_context_solve(ctx, name, build, kw):
    prob = build()
    prm = params(**kw)
    v0, i0, inf0 = context_solve(ctx, prob, prm)
    v1, i1, inf1 = noalloc_solve(prob, prm)
    assert inf1.iters == inf0.iters
    assert np.array_equal(v1, v0) and np.array_equal(i1, i0)


def test_device_gyro_and_no_warm_start(ctx):
    # rigid bodies with unequal principal inertias, spinning: f_gyro computed on the device
    prob = scenes.random_problem(nb=400, nc=600, seed=11, vscale=3.0)
    prm = params(tol_abs=1e-9, tol_rel=1e-9)
    v0, i0, inf0 = context_solve(ctx, prob, prm, with_fgyro=False, with_warm=False)
    v1, i1, inf1 = noalloc_solve(prob, prm, with_fgyro=False, with_warm=False)
    assert inf1.iters == inf0.iters
    assert np.array_equal(v1, v0) and np.array_equal(i1, i0)


def test_no_contacts_is_the_free_solve(ctx):
    prob = scenes.random_problem(nb=50, nc=0, seed=1)
    v0, _, inf0 = context_solve(ctx, prob, params())
    v1, _, inf1 = noalloc_solve(prob, params())
    assert inf0.iters == inf1.iters == 0
    assert np.array_equal(v1, v0)


def test_errors():
    prob = scenes.random_problem(nb=30, nc=40, seed=2)
    # beyond the static maximum: refused before any device work
    with pytest.raises(native.CdxError) as e:
        native.pj_solve_noalloc(20000, 10, {}, 1e-3, 1.0, params(), 0, 0)
    assert e.value.status == native.CDX_BAD_ARG and "maximum" in str(e.value)
    with pytest.raises(native.CdxError) as e:
        native.pj_solve_noalloc(10, 70000, {}, 1e-3, 1.0, params(), 0, 0)
    assert e.value.status == native.CDX_BAD_ARG
    with pytest.raises(native.CdxError) as e:
        noalloc_solve(prob, params(accel=native.CDX_ACCEL_NESTEROV))
    assert e.value.status == native.CDX_BAD_ARG
    bad = prob.copy()
    bad.body_b = bad.body_b.copy()
    bad.body_b[3] = prob.nb + 5  # out of range
    with pytest.raises(native.CdxError) as e:
        noalloc_solve(bad, params())
    assert e.value.status == native.CDX_BAD_ARG
    bad = prob.copy()
    bad.frame = bad.frame.copy()
    bad.frame[2, 4] = np.inf
    with pytest.raises(native.CdxError) as e:
        noalloc_solve(bad, params())
    assert e.value.status == native.CDX_NONFINITE
    bad = prob.copy()
    bad.minv = bad.minv.copy()
    bad.minv[int(bad.dof[:2].sum()) + 0:int(bad.dof[:2].sum()) + 3] = 0.0  # zero inverse mass on a dynamic body
    with pytest.raises(native.CdxError) as e:
        noalloc_solve(bad, params())
    assert e.value.status == native.CDX_NOT_FACTORIZED
    bad = prob.copy()
    bad.dof = bad.dof.copy()
    bad.dof[0] = 5
    with pytest.raises(native.CdxError) as e:
        noalloc_solve(bad, params())
    assert e.value.status == native.CDX_BAD_ARG
    # a good call after the failures is unaffected
    v1, i1, inf1 = noalloc_solve(prob, params(tol_abs=1e-8, tol_rel=1e-8))
    assert np.isfinite(v1).all() and inf1.iters > 0
