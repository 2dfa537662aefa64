"""The oracle's discrete adjoint (oracle/lassie_oracle_adjoint.c), pinned
without reference to its own formulas:
  - replaying a recorded step sequence reproduces the solve bit for bit;
  - closed forms for y' = -k y: the RKF45 map is the Fehlberg polynomial
    P(z) (z = -k dt) built here from the tableau, so d log y_N / d log k =
    sum z P'(z) / P(z); a BDF1 step is y / (1 + h k);
  - reverse mode (adjoint) equals forward mode (tangent) on every parameter,
    two independent code paths (transposed vs forward solves, vector-Jacobian
    vs Jacobian-vector products);
  - the tangent equals central differences of the replay, including models
    whose Newton Jacobian differs from df/dx (createJac_ drops the exponents
    of the other reactants, kernels.cuh:248-249) and models with feed species.
"""
from fractions import Fraction as F

import numpy as np
import pytest
from numpy.polynomial import Polynomial as Poly

from lassie_amd import synth
from oracle import oracle as O
from tests import cases

CTRL = dict(rtol=1e-6, bdf_regrow=1, max_rkf_retries=8)


def _fehlberg_w1_poly():
    z = Poly([0, 1])
    one = Poly([1])
    k1 = z
    k2 = z * (one + k1 / 4)
    k3 = z * (one + float(F(3, 32)) * k1 + float(F(9, 32)) * k2)
    k4 = z * (one + float(F(1932, 2197)) * k1 - float(F(7200, 2197)) * k2 + float(F(7296, 2197)) * k3)
    k5 = z * (one + float(F(439, 216)) * k1 - 8 * k2 + float(F(3680, 513)) * k3 - float(F(845, 4104)) * k4)
    return one + float(F(25, 216)) * k1 + float(F(1408, 2565)) * k3 + float(F(2197, 4104)) * k4 - 0.2 * k5


def _kinds(tape):
    k, _ = tape.steps()
    return set(np.unique(k).tolist())


def _dirderiv(m, tape, nt, W, d_logk, d_x0=None):
    return float(np.sum(W * O.tangent(m, tape, nt, d_logk, d_x0)))


@pytest.mark.parametrize("model,t,kw", [
    (cases.mixed(), synth.t_vector(0, 3, 7), {}),
    (cases.with_feed(), synth.t_vector(0, 2, 5), {}),
    (synth.stiff(24, 30, 2), synth.t_vector(0, 1, 6), CTRL),
])
def test_replay_reproduces_the_solve(model, t, kw):
    out, st, rc, tape = O.solve_tape(model, t, **kw)
    assert rc == 0 and tape.n == st["rkf_accepts"] + st["bdf_steps"]
    assert np.array_equal(O.replay(model, tape, len(t)), out)


def test_rkf_decay_closed_form():
    k = 0.9
    m = cases.decay(k=k, x0=1.0)
    t = synth.t_vector(0, 2, 5)
    out, _, _, tape = O.solve_tape(m, t)
    kinds, dts = tape.steps()
    assert _kinds(tape) == {O.STEP_RKF}
    P = _fehlberg_w1_poly()
    dP = P.deriv()
    W = np.zeros_like(out)
    W[-1, 0] = 1.0
    gk, gx = O.adjoint(m, tape, W)
    z = -k * dts
    expect = out[-1, 0] * np.sum(z * dP(z) / P(z))
    assert gk[0] == pytest.approx(expect, rel=1e-12)
    assert gx[0] == pytest.approx(np.prod(P(z)), rel=1e-12)


def test_bdf_decay_closed_form():
    k, h, n = 3.0, 0.01, 25
    m = cases.decay(k=k)
    m.atol = np.zeros(1)  # every RKF attempt rejects with factor 0: BDF1 steps of be_step
    t = np.arange(n + 1) * h
    out, _, rc, tape = O.solve_tape(m, t, be_step=h, max_iters=10 * n)
    kinds, _ = tape.steps()
    assert rc == 0 and np.all(kinds >= 1)
    W = np.zeros_like(out)
    W[-1, 0] = 1.0
    gk, gx = O.adjoint(m, tape, W)
    assert gk[0] == pytest.approx(-n * h * k / (1 + h * k) * out[-1, 0], rel=1e-10)
    assert gx[0] == pytest.approx((1 + h * k) ** -n, rel=1e-10)


def _quirky():
    """2A + B -> C and C -> A: createJac_'s Jacobian differs from df/dx."""
    return cases.from_reactions(3, [({0: 2, 1: 1}, {2: 1}), ({2: 1}, {0: 1})], [2.5, 0.3], [1.0, 0.8, 0.1])


def _case(name):
    """Models whose step sequences mix RKF45 and BDF1 steps (a tighter atol
    sends the small hand-built models through BDF1 as well)."""
    if name == "mixed":
        m, t, kw = cases.mixed(), synth.t_vector(0, 3, 7), {}
    elif name == "feed":
        m, t, kw = cases.with_feed(), synth.t_vector(0, 2, 5), {}
    elif name == "quirky":
        m, t, kw = _quirky(), synth.t_vector(0, 2, 5), {}
    else:
        return synth.stiff(16, 24, 7), synth.t_vector(0, 1, 6), CTRL
    m.atol = np.full(m.S, 1e-14)
    return m, t, kw


@pytest.mark.parametrize("name", ["mixed", "feed", "quirky", "stiff"])
def test_adjoint_equals_tangent_on_every_parameter(name):
    m, t, kw = _case(name)
    out, _, rc, tape = O.solve_tape(m, t, **kw)
    assert rc == 0
    assert any(kd >= 1 for kd in _kinds(tape)), "the sequence must contain BDF1 steps"
    rng = np.random.default_rng(5)
    W = rng.standard_normal(out.shape)
    gk, gx = O.adjoint(m, tape, W)
    scale = max(np.abs(gk).max(), 1e-300)
    for r in range(m.R):
        e = np.zeros(m.R)
        e[r] = 1.0
        assert abs(gk[r] - _dirderiv(m, tape, len(t), W, e)) <= 1e-10 * scale, r
    dx = rng.standard_normal(m.S)
    assert gx @ dx == pytest.approx(_dirderiv(m, tape, len(t), W, np.zeros(m.R), dx), rel=1e-9, abs=1e-12)
    if name == "feed":
        fed = m.feed != 0
        # after the first step the fed species is pinned; only sample 0 sees its x0
        W0 = W.copy()
        W0[0] = 0.0
        assert np.all(O.adjoint(m, tape, W0)[1][fed] == 0.0)


@pytest.mark.parametrize("name", ["mixed", "quirky", "stiff"])
def test_tangent_matches_central_differences(name):
    m, t, kw = _case(name)
    out, _, _, tape = O.solve_tape(m, t, **kw)
    rng = np.random.default_rng(9)
    W = rng.standard_normal(out.shape)
    logk = np.log(m.k)
    h = 1e-4
    for r in range(min(m.R, 8)):
        e = np.zeros(m.R)
        e[r] = 1.0
        fp = np.sum(W * O.replay(m, tape, len(t), k=np.exp(logk + h * e)))
        fm = np.sum(W * O.replay(m, tape, len(t), k=np.exp(logk - h * e)))
        fd = (fp - fm) / (2 * h)
        tl = _dirderiv(m, tape, len(t), W, e)
        assert tl == pytest.approx(fd, rel=1e-6, abs=1e-9), r


def test_log_loss_gradient_within_1e5_of_fd_h1e6():
    """The differentiation form's loss, sum (log(y+eps) - log(y*+eps))^2 over
    the samples, checked with central differences of step h = 1e-6 in log k:
    |adjoint - fd| <= 1e-5 * max |gradient|."""
    m = synth.random_model(20, 40, 11, k_range=(1e-4, 1e4), x0_kind="loguniform", x0_range=(1e-6, 1.0))
    t = synth.t_vector(0, 1, 6)
    out, _, rc, tape = O.solve_tape(m, t, **CTRL)
    assert rc == 0
    eps = 1e-12
    rng = np.random.default_rng(3)
    target = (out + eps) * np.exp(0.2 * rng.standard_normal(out.shape))

    def loss(o):
        return float(np.sum((np.log(o + eps) - np.log(target)) ** 2))

    gout = 2.0 * (np.log(out + eps) - np.log(target)) / (out + eps)
    gk, _ = O.adjoint(m, tape, gout)
    logk = np.log(m.k)
    h = 1e-6
    scale = np.abs(gk).max()
    for r in range(m.R):
        e = np.zeros(m.R)
        e[r] = h
        fd = (loss(O.replay(m, tape, len(t), k=np.exp(logk + e)))
              - loss(O.replay(m, tape, len(t), k=np.exp(logk - e)))) / (2 * h)
        assert abs(gk[r] - fd) <= 1e-5 * scale, (r, gk[r], fd)


def test_adjoint_overflows_only_through_a_nonconverged_newton_step():
    """Parameterisation 108 of the 512 x 1024 differentiation workload
    (k logU[1e-4, 1e4], t in [0, 0.2]) starts with a BDF1 step whose modified
    Newton iteration runs all 10000 iterations without converging: the last
    correction's norm is still about 1 (the loop keeps that iterate,
    Simulation.cuh:789, 883-890).  The derivative of that iterate grows
    geometrically with the iterations, so the exact adjoint and the tangent
    overflow; parameterisation 0, whose Newton steps converge, has a finite
    gradient.  The GPU reports the same split (LASSIE_ERR_NONFINITE)."""
    from lassie_amd import synth
    models, _ = synth.parameterisations(512, 1024, 109, seed=0)
    t = synth.t_vector(0.0, 0.2, 50)
    ctrl = dict(rtol=1e-6, bdf_regrow=1, max_rkf_retries=8)
    rng = np.random.default_rng(1)
    for b, finite in ((0, True), (108, False)):
        m = models[b]
        out, st, rc, tape = O.solve_tape(m, t, **ctrl)
        kind, dt = tape.steps()
        assert rc == 0 and (kind.max() == 10000) == (not finite)
        gk, gx = O.adjoint(m, tape, rng.standard_normal(out.shape))
        assert np.isfinite(gk).all() == finite and np.isfinite(gx).all() == finite
        dout = O.tangent(m, tape, len(t), rng.standard_normal(m.R))
        assert np.isfinite(dout).all() == finite
    # the non-converged step: the first one, BDF1 with step dt[0] from x0
    m = models[108]
    _, _, _, tape = O.solve_tape(m, t, **ctrl)
    kind, dt = tape.steps()
    assert kind[0] == 10000
    y = np.array(m.x0, float)
    LU, piv, info = O.lu(O.newton_matrix(O.jac(m, y), dt[0]))
    assert info == 0
    v = y.copy()
    for _ in range(1000):
        x = O.lu_solve(LU, piv, -(v - y - dt[0] * O.rhs(m, v)))
        v = v + x
    assert np.sqrt(O.normsq(x)) > 1e-2  # far from newton_tol = 1e-6 after 1000 of the 10000 iterations
