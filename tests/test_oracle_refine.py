"""The oracle's precision form: FP32 LU of (float)M with FP64 iterative
refinement (orc_refine_solve) inside the FP64 integrator.

Pinned by: exact rational residuals (fractions) of the refined solution
against the Oettli-Prager bound; numpy's FP64 solve; the first sweep equal
to the FP32 solve of (float)rhs; a target no sweep can meet (tol = 0) and an
ill-conditioned matrix (Hilbert, kappa ~ 1e10 > 1/eps_32), which must not
converge; a component under FP32's range, admitted only by the absolute
floor; and, for whole solves, the structural fact that with tol = 0 and
one sweep every BDF step falls back to the FP64 LU, so the trajectory equals
the FP64 form bit for bit.
"""
from fractions import Fraction

import numpy as np
import pytest

from lassie_amd import synth
from oracle import oracle as O


def _factor32(M):
    LU32, piv, info = O.lu32(M.astype(np.float32))
    assert info == 0
    return LU32, piv


def _exact_residual(M, x, b):
    n = M.shape[0]
    r = []
    for i in range(n):
        acc = Fraction(b[i])
        for j in range(n):
            acc -= Fraction(M[i, j]) * Fraction(x[j])
        r.append(acc)
    return r


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_refined_solution_meets_the_componentwise_bound(seed):
    rng = np.random.default_rng(seed)
    n = 24
    M = rng.uniform(-1, 1, (n, n)) + n * np.eye(n)
    M[rng.uniform(size=(n, n)) < 0.3] = 0.0
    np.fill_diagonal(M, n + rng.uniform(0, 1, n))
    b = rng.uniform(-1, 1, n) * 10.0 ** rng.integers(-6, 3, n)
    LU32, piv = _factor32(M)
    tol = 1e-14
    x, r, ok, sweeps = O.refine_solve(M, LU32, piv, b, tol=tol, max_it=4)
    assert ok and 1 <= sweeps <= 4
    np.testing.assert_allclose(x, np.linalg.solve(M, b), rtol=1e-12, atol=0)
    exact = _exact_residual(M, x, b)
    mag = np.abs(b) + np.abs(M) @ np.abs(x)
    for i in range(n):
        # the computed residual carries at most ~n eps |M||x| of rounding
        assert abs(float(exact[i])) <= 1.01 * tol * mag[i] + 2 * n * 2.0 ** -53 * mag[i]
        assert abs(float(exact[i]) - r[i]) <= 2 * n * 2.0 ** -53 * mag[i]


def test_first_sweep_is_the_fp32_solve():
    rng = np.random.default_rng(5)
    n = 16
    M = rng.uniform(-1, 1, (n, n)) + 4 * np.eye(n)
    b = rng.uniform(-1, 1, n)
    LU32, piv = _factor32(M)
    x, _, _, sweeps = O.refine_solve(M, LU32, piv, b, tol=0.0, max_it=1)
    assert sweeps == 1
    d = O.lu_solve32(LU32, piv, b.astype(np.float32))
    np.testing.assert_array_equal(x, d.astype(np.float64))


def test_unreachable_targets_do_not_converge():
    rng = np.random.default_rng(6)
    n = 12
    M = rng.uniform(-1, 1, (n, n)) + 3 * np.eye(n)
    b = rng.uniform(-1, 1, n)
    LU32, piv = _factor32(M)
    _, _, ok, sweeps = O.refine_solve(M, LU32, piv, b, tol=0.0, max_it=3)
    assert not ok and sweeps == 3
    H = np.array([[1.0 / (i + j + 1) for j in range(9)] for i in range(9)])
    LU32, piv = _factor32(H)
    _, _, ok, sweeps = O.refine_solve(H, LU32, piv, np.ones(9), tol=1e-14, max_it=4)
    assert not ok and sweeps == 4


def test_absolute_floor_admits_components_below_fp32_range():
    """x_1 = -1e-78 x_0 / 2 underflows in FP32, so no sweep can reach a
    componentwise target for it; with the species' absolute tolerance as
    floor (1e-12) the first sweep meets the target."""
    M = np.array([[1.0, 0.0], [1e-78, 2.0]])
    b = np.array([1.0, 0.0])
    LU32, piv = _factor32(M)
    x, _, ok, sweeps = O.refine_solve(M, LU32, piv, b, tol=1e-14, max_it=4)
    assert not ok and sweeps == 4 and x[1] == 0.0
    x, r, ok, sweeps = O.refine_solve(M, LU32, piv, b, tol=1e-14, max_it=4, floor=np.full(2, 1e-12))
    assert ok and sweeps == 1 and x[0] == 1.0
    assert abs(r[1]) == 1e-78


BENCH = dict(rtol=1e-6, bdf_regrow=1, max_rkf_retries=8)
COUNTERS = ("loop_iters", "rkf_attempts", "rkf_accepts", "bdf_steps", "newton_iters",
            "lu_failures", "nan_repairs", "samples")


@pytest.mark.parametrize("S,seed", [(8, 11), (24, 12), (40, 13)])
def test_precision_form_equals_fp64_when_every_step_falls_back(S, seed):
    m = synth.stiff(S, 2 * S, seed, (1e-4, 1e4))
    t = synth.t_vector(0, 2, 5)
    a, sa, _, _, ya = O.solve(m, t, **BENCH)
    b, sb, _, _, yb = O.solve(m, t, refine=1, refine_tol=0.0, refine_max_iter=1, **BENCH)
    assert sa["bdf_steps"] > 0
    for k in COUNTERS:
        assert sa[k] == sb[k], k
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(ya, yb)
    # one FP32 sweep, then the FP64 LU, in every BDF step that factors
    steps = sb["bdf_steps"] - sb["lu_failures"]
    assert sb["refine_fallbacks"] == steps and sb["refine_iters"] == steps


@pytest.mark.parametrize("S,seed", [(48, 3), (64, 4), (32, 5)])
def test_precision_form_converges_on_mild_steps(S, seed):
    """Stiff models with BDF steps of 1e-3: refinement meets the default
    1e-14 componentwise target without falling back, the controller takes
    the same decisions as in the FP64 form, and the trajectory stays within
    1e-5 of it (|a-b| / (|b| + atol), atol 1e-12; two of the three models
    give identical trajectories, the third drifts to 4.9e-6 through the
    controller's step-size feedback)."""
    m = synth.stiff(S, 2 * S, seed, (1e-4, 1e2))
    t = synth.t_vector(0, 1, 11)
    kw = dict(be_step=1e-3, **BENCH)
    a, sa, *_ = O.solve(m, t, **kw)
    b, sb, *_ = O.solve(m, t, refine=1, refine_tol=1e-14, refine_max_iter=6, **kw)
    assert sb["bdf_steps"] > 0 and sb["refine_iters"] >= sb["newton_iters"]
    assert sb["refine_fallbacks"] == 0
    for k in COUNTERS:
        assert sa[k] == sb[k], k
    rel = np.abs(b - a) / (np.abs(a) + 1e-12)
    assert rel.max() <= 1e-5


def test_zero_sweeps_is_the_fp32_solve():
    """max_it = 0: the FP32 solve of (float)rhs alone, no residual and no
    test, reported as converged after one solve."""
    rng = np.random.default_rng(7)
    n = 20
    M = rng.uniform(-1, 1, (n, n)) + 4 * np.eye(n)
    b = rng.uniform(-1, 1, n) * 10.0 ** rng.integers(-8, 3, n)
    LU32, piv = _factor32(M)
    x, _, ok, sweeps = O.refine_solve(M, LU32, piv, b, tol=0.0, max_it=0)
    assert ok and sweeps == 1
    d = O.lu_solve32(LU32, piv, b.astype(np.float32))
    np.testing.assert_array_equal(x, d.astype(np.float64))
    # the FP32 solve alone is about eps_32 * kappa from the FP64 solution,
    # far from what one refinement sweep reaches
    rel = np.abs(x - np.linalg.solve(M, b)) / np.abs(np.linalg.solve(M, b))
    assert 1e-9 < rel.max() < 1e-5


@pytest.mark.parametrize("S,seed,tend", [(24, 12, 2.0), (48, 3, 1.0), (100, 1001, 5.0)])
def test_unrefined_form_is_refinement_with_an_unbounded_target(S, seed, tend):
    """Whole solves: refine_max_iter = 0 (one FP32 solve per Newton
    iteration) equals the refined form with a target every first sweep
    meets (refine_tol = 1e300), bit for bit, counters included: both take
    x = 0 + (double)(FP32 solve) as the Newton correction."""
    m = synth.stiff(S, 2 * S, seed, (1e-4, 1e4))
    t = synth.t_vector(0, tend, 11)
    a, sa, _, _, ya = O.solve(m, t, refine=1, refine_max_iter=0, **BENCH)
    b, sb, _, _, yb = O.solve(m, t, refine=1, refine_tol=1e300, refine_max_iter=4, **BENCH)
    assert sa["bdf_steps"] > 0
    for k in COUNTERS + ("refine_iters", "refine_fallbacks"):
        assert sa[k] == sb[k], k
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(ya, yb)
    # one FP32 solve per Newton iteration; fallbacks only at zero FP32 pivots
    assert sa["refine_iters"] == sa["newton_iters"] and sa["refine_fallbacks"] == 0
