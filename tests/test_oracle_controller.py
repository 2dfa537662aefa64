"""Hand-computed single-step traces of the oracle's step controller on one-
species models, one test per rule of Simulation<T>::run
(src/c++/header/Simulation.cuh:670-919):

  Q1  on an accepted RKF45 step, y1 <- w1 (computed with the step just taken)
      but t advances by the *next* step dt*min(delta) (:753-767);
  clamp  that next step is cut to t1 - t when t + dt*min(delta) > t1 (:759-764),
      and a BDF1 step is cut the same way (:783-786);
  sampling  a sample is stored when |t1 - t| < 1e-5, absolute (:674-676);
  Q3  after a rejection the loop switches to BDF1 when dt*min(delta) <
      stiffness_tol, is NaN/inf, or |dt_old - dt| < 1e-3, absolute (:774-778);
  Newton  modified Newton stops when sqrt(sum float(d_s^2)) <= newton_tol,
      an absolute norm (:789, 883-890).

The observables are the progress log (t / t_max at the top of every loop
iteration, the ProgressSimulation file, :671), the stored samples, the final
state and the counters.  The expected values come from the rule as written
and from one RKF45 attempt of the oracle (O.rkf_attempt, itself pinned by the
exact rational Fehlberg tableau in test_oracle.py); each case is chosen so
that the plausible mistake changes the outcome: advancing t by the step taken,
not clamping, a relative sampling or switch test, a relative Newton norm.
"""
import numpy as np

from oracle import oracle as O
from tests import cases


def _attempt(m, y, dt, rtol=0.0):
    nok, w1, w2, delta = O.rkf_attempt(m, np.array(y, float), dt, rtol)
    return nok, w1, O.min_first(delta)


def test_accept_advances_time_by_the_next_step():
    """One accepted step of A -> 0 (k = 2, x0 = 1e-3) from rkf_step = 1e-3: the state
    becomes w1 of the step taken (dt = 1e-3), but t advances by the *next*
    step 1e-3 * mind, which is also the step of the following attempt.
    t += dt_old would give t = 1e-3."""
    m = cases.decay(k=2.0, x0=1e-3)
    tv = np.array([0.0, 10.0])
    nok0, w1a, mind0 = _attempt(m, [1e-3], 1e-3)
    # accepted, with a finite next step that differs from the taken one
    assert nok0 == 1 and 2.0 < mind0 < 100.0
    dt1 = 1e-3 * mind0
    _, st, rc, prog, yf = O.solve(m, tv, progress_cap=2, max_iters=2)
    np.testing.assert_array_equal(prog, np.array([0.0, dt1]) / 10.0)
    _, st1, _, _, yf1 = O.solve(m, tv, max_iters=1)
    np.testing.assert_array_equal(yf1, w1a)
    assert st1["rkf_accepts"] == 1 and st1["t_final"] == dt1 and st1["dt_final"] == dt1
    # the second attempt is made with dt1 from w1a: its outcome decides the state
    nok1, w1b, mind1 = _attempt(m, w1a, dt1)
    if nok1 == 1:
        np.testing.assert_array_equal(yf, w1b)
        assert st["t_final"] == dt1 + dt1 * mind1
    else:
        np.testing.assert_array_equal(yf, w1a)
        assert st["t_final"] == dt1 and st["dt_final"] == dt1 * mind1


def test_accept_clamps_the_next_step_to_the_sample_time():
    """Samples at 0 and 2e-3: the first step (dt = 1e-3) is accepted and its
    next step 1e-3 * mind > 2e-3 is cut to t1 - t = 2e-3 (t is still 0 when
    the test is made), so t lands on 2e-3 and the next iteration stores the
    state w1(1e-3) as sample 1 (the state is one step behind its label)."""
    m = cases.decay(k=2.0, x0=1e-3)
    tv = np.array([0.0, 2e-3, 1.0])
    nok, w1a, mind = _attempt(m, [1e-3], 1e-3)
    assert nok == 1 and np.isfinite(mind) and 1e-3 * mind > 2e-3
    out, st, rc, prog, yf = O.solve(m, tv, progress_cap=2, max_iters=2)
    np.testing.assert_array_equal(prog, np.array([0.0, 2e-3]) / 1.0)
    assert st["samples"] == 2
    assert out[1, 0] == w1a[0]
    # the attempt after the sample uses the clamped step
    nok2, w1b, _ = _attempt(m, w1a, 2e-3)
    assert nok2 == 1
    np.testing.assert_array_equal(yf, w1b)


def test_sampling_tolerance_is_absolute_1e_minus_5():
    """At t = 0 a first sampling time of 9.9e-6 is within 1e-5 and is stored
    at once (the initial state); 1.01e-5 is not, so the loop first takes the
    step dt = 1e-3, cut to 1.01e-5 for t, and stores w1(1e-3) at the next
    iteration.  A test relative to t1 would store neither at t = 0."""
    m = cases.decay(k=2.0, x0=1e-3)
    out, st, _, _, _ = O.solve(m, np.array([9.9e-6, 1.0]), max_iters=1)
    assert st["samples"] == 1 and out[0, 0] == 1e-3
    out, st, _, _, _ = O.solve(m, np.array([1.01e-5, 1.0]), max_iters=1)
    assert st["samples"] == 0
    _, w1a, _ = _attempt(m, [1e-3], 1e-3)
    out, st, _, prog, _ = O.solve(m, np.array([1.01e-5, 1.0]), max_iters=2, progress_cap=2)
    assert st["samples"] == 1 and out[0, 0] == w1a[0]
    assert prog[1] == 1.01e-5


def test_stiffness_switch_is_absolute_1e_minus_3():
    """A stiff decay (k = 100) rejects its first attempt.  From rkf_step =
    1e-3 the shrunken step stays above stiffness_tol but differs from 1e-3 by
    less than 1e-3, so the loop switches to BDF1 at once (a relative 1e-3
    test would retry RKF45).  From rkf_step = 0.5 the difference is far above
    1e-3 and the loop retries RKF45 with the shrunken step."""
    m = cases.decay(k=100.0, x0=1.0)
    tv = np.array([0.0, 10.0])
    nok, _, mind = _attempt(m, [1.0], 1e-3)
    assert nok == 0 and 1e-3 * mind > 1e-6 and abs(1e-3 - 1e-3 * mind) < 1e-3
    assert abs(1e-3 - 1e-3 * mind) > 1e-3 * 1e-3
    _, st, _, _, _ = O.solve(m, tv, max_iters=1)
    assert st["rkf_attempts"] == 1 and st["bdf_steps"] == 1
    assert st["t_final"] == 0.01  # the BDF1 step is newDtBDF = be_step, and t advances by it
    nok, _, mind = _attempt(m, [1.0], 0.5)
    assert nok == 0 and abs(0.5 - 0.5 * mind) >= 1e-3 and 0.5 * mind > 1e-6
    _, st, _, _, _ = O.solve(m, tv, max_iters=1, rkf_step=0.5)
    assert st["rkf_attempts"] == 1 and st["bdf_steps"] == 0
    assert st["dt_final"] == 0.5 * mind  # retried with the shrunken step


def test_bdf_step_is_cut_to_the_sample_time():
    """A BDF1 step that would pass t1 is cut to t1 - t: with samples at 0 and
    4e-3 the stiff decay takes one backward-Euler step of 4e-3, stored as
    sample 1: x0 / (1 + 4e-3 k) up to the Newton iteration's rounding."""
    k = 1e4
    m = cases.decay(k=k, x0=1.0)
    tv = np.array([0.0, 4e-3, 1.0])
    _, st, _, _, _ = O.solve(m, tv, max_iters=1)
    assert st["bdf_steps"] == 1 and st["t_final"] == 4e-3 and st["dt_final"] == 4e-3
    out, st, _, prog, _ = O.solve(m, tv, max_iters=2, progress_cap=2)
    assert st["samples"] == 2 and prog[1] == 4e-3
    assert abs(out[1, 0] - 1.0 / (1.0 + 4e-3 * k)) <= 4 * np.finfo(float).eps * out[1, 0]


def test_newton_stops_on_the_absolute_norm():
    """Backward Euler on a linear decay is solved exactly by the first
    Newton correction d1 = -hk y/(1 + hk).  From y = 1, |d1| ~ 0.99 exceeds
    1e-6 and a second iteration confirms (2 iterations); from y = 1e-9,
    |d1| ~ 1e-9 is already below the absolute tolerance (1 iteration).  A
    norm relative to y would take 2 iterations in both cases."""
    tv = np.array([0.0, 10.0])
    for x0, iters in ((1.0, 2), (1e-9, 1)):
        m = cases.decay(k=1e4, x0=x0)
        _, st, _, _, yf = O.solve(m, tv, max_iters=1)
        assert st["bdf_steps"] == 1 and st["newton_iters"] == iters, (x0, st["newton_iters"])
        h = 0.01
        # one correction from y1 = x0 is exact up to rounding at the scale of x0
        assert abs(yf[0] - x0 / (1.0 + h * 1e4)) <= 4 * np.finfo(float).eps * x0
    # the same step with newton_tol above |d1| of y = 1 stops after one iteration
    m = cases.decay(k=1e4, x0=1.0)
    _, st, _, _, _ = O.solve(m, tv, max_iters=1, newton_tol=1.0)
    assert st["bdf_steps"] == 1 and st["newton_iters"] == 1
