"""GPU: the state transfers and the device finiteness count.  cdx_set_contacts checks the
contacts over contiguous chunks in host threads (from 32,768 contacts) and stages them in pinned
memory; the first failing contact still names the error, whichever chunk it lies in.
cdx_get_velocities / cdx_get_positions / cdx_get_impulses pack the stacked layout (sum(dof) twists, 7 or 3 pose entries
per body, [nc][3] impulses) on the device and copy it through a pinned staging buffer that grows
with the problem; cdx_nonfinite_bodies counts the bodies whose solved twist or pose holds a NaN or
Inf.  Mixed rigid bodies and point masses; a context that solves a large, a small and a large
problem again gives the bits of fresh contexts, and the count equals the count of the downloaded
state."""
import numpy as np
import pytest

import oracle
from cardillocxx_amd import native, scenes

pytestmark = pytest.mark.gpu


def params(**kw):
    p = native.default_params()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def problem(nb, nc, seed):
    return scenes.random_problem(nb=nb, nc=nc, seed=seed, p_point=0.3)


def run(c, prob, k=30):
    c.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
    q0 = c.positions()
    c.set_contacts(*prob.contact_arrays())
    c.pj_solve(1e-3, 1.0, params(max_iter=k, tol_abs=1e-15, tol_rel=1e-15))
    return q0, c.velocities(), c.impulses()


def per_body_bad(dof, v, q):
    """Bodies with a non-finite entry in their stacked twist or pose."""
    nq = np.where(dof == 6, 7, 3)
    vo = np.concatenate([[0], np.cumsum(dof)])
    qo = np.concatenate([[0], np.cumsum(nq)])
    return np.array([not (np.isfinite(v[vo[b]:vo[b + 1]]).all() and np.isfinite(q[qo[b]:qo[b + 1]]).all())
                     for b in range(len(dof))])


def test_reused_context_gives_fresh_bits():
    probs = [problem(400, 1200, 1), problem(30, 70, 2), problem(500, 1500, 3)]
    reused = native.Context(0)
    try:
        for prob in probs:
            q0, v, imp = run(reused, prob)
            fresh = native.Context(0)
            try:
                f0, fv, fimp = run(fresh, prob)
            finally:
                fresh.close()
            np.testing.assert_array_equal(q0, prob.q)  # positions as set, 7 or 3 entries per body
            np.testing.assert_array_equal(q0, f0)
            np.testing.assert_array_equal(v, fv)
            np.testing.assert_array_equal(imp, fimp)
            st, ov, oimp, res = oracle.pj_solve(prob, 1e-3, oracle.params(max_iter=30, tol_abs=1e-15, tol_rel=1e-15))
            assert st == 0
            scale = max(np.abs(ov).max(), 1e-300)
            assert np.abs(v - ov).max() / scale <= 1e-10
            assert np.abs(imp - oimp).max() / max(np.abs(oimp).max(), 1e-300) <= 1e-10
            assert v.shape == (prob.nv,) and imp.shape == (prob.nc, 3)
    finally:
        reused.close()


def test_nonfinite_count_matches_the_downloaded_state():
    prob = problem(600, 1800, 5)
    c = native.Context(0)
    try:
        c.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
        with pytest.raises(native.CdxError):
            c.nonfinite_bodies()  # no solve yet
        c.set_contacts(*prob.contact_arrays())
        c.pj_solve(1e-3, 1.0, params(max_iter=20))
        assert c.nonfinite_bodies() == 0
        # a step of 1e308 s overflows the positions of the bodies moving faster than ~1.8 m/s or rad/s
        # (a non-finite quaternion normalises to NaN); the slower ones stay finite
        c.moreau_update(1e308, 1.0)
        v, q = c.velocities(), c.positions()
        bad = per_body_bad(prob.dof, v, q)
        assert 0 < bad.sum() < len(bad)
        assert c.nonfinite_bodies() == int(bad.sum())
    finally:
        c.close()


def test_threaded_contact_checks_name_the_first_failure():
    # ten copies of a 7,000-contact problem side by side (70,000 contacts: two host threads' chunks)
    prob = scenes.concat_problems([problem(2000, 7000, 7)] * 10)[0]
    assert prob.nc == 70000
    c = native.Context(0)
    try:
        c.set_bodies(prob.dof, prob.minv, prob.v, prob.fext, prob.fgyro, prob.q)
        base = [np.array(a, copy=True) for a in prob.contact_arrays()]
        cases = [
            # (contact, array index, entry, value) planted in order; the lowest contact's check fails
            ([(60000, 3, 2, np.nan), (65000, 0, None, prob.nb + 5)], "non-finite contact frame"),
            ([(100, 0, None, prob.nb + 5), (50000, 2, None, np.inf)], "body index out of range"),
            ([(69999, 8, 1, np.inf), (40000, 2, None, np.nan)], "non-finite mu"),
            ([(1234, 6, 0, np.inf)], "non-finite contact data"),
        ]
        for plants, msg in cases:
            arrs = [a.copy() for a in base]
            for cc, k, e, val in plants:
                if e is None:
                    arrs[k][cc] = val
                else:
                    arrs[k][cc, e] = val
            with pytest.raises(native.CdxError) as err:
                c.set_contacts(*arrs)
            assert msg in str(err.value)
        # the rejected calls leave a context that takes the valid contacts and solves them
        c.set_contacts(*base)
        c.pj_solve(1e-3, 1.0, params(max_iter=30, tol_abs=1e-15, tol_rel=1e-15))
        st, ov, oimp, res = oracle.pj_solve(prob, 1e-3, oracle.params(max_iter=30, tol_abs=1e-15, tol_rel=1e-15))
        assert st == 0
        assert np.abs(c.velocities() - ov).max() / np.abs(ov).max() <= 1e-10
        assert np.abs(c.impulses() - oimp).max() / np.abs(oimp).max() <= 1e-10
    finally:
        c.close()
