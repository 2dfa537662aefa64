"""LSODAsim (lassie_amd/lsodasim.py): the reference's CPU LSODA simulator.

Checked against the oracle's independent LSODA restatement (oracle/lsoda.py)
on order <= 2 models, closed forms for the generator's rules (a reactant
stoichiometry of 3 contributes no factor; feed species are frozen at their
feed value), the t = 0 prepend rule, the fitness mode, the stochastic
conversion, and central differences for the Jacobian.
"""
import numpy as np
import pytest

from lassie_amd import lsodasim, synth
from oracle import lsoda
from tests import cases


def _write(tmp_path, m, t, **kw):
    d = tmp_path / "in"
    synth.write_folder(m, str(d), t, **kw)
    return d


def _run(d, out, *flags):
    assert lsodasim.main([str(d), str(out), *flags]) == 0
    return np.loadtxt(out / "ODESol", ndmin=2)


@pytest.mark.parametrize("name", ["mixed", "chain", "reversible", "dimer"])
def test_odesol_matches_the_oracle_lsoda(tmp_path, name):
    m = getattr(cases, name)()
    t = np.linspace(0.0, 3.0, 7)
    sol = _run(_write(tmp_path, m, t), tmp_path / "out")
    assert sol.shape == (7, m.S + 1)
    np.testing.assert_array_equal(sol[:, 0], np.loadtxt(tmp_path / "in" / "t_vector"))
    ref = lsoda.solve(m, t)
    np.testing.assert_allclose(sol[:, 1:], ref, rtol=2e-5, atol=1e-10)
    text = (tmp_path / "out" / "ODESol").read_text().splitlines()[1].split("\t")
    assert all(len(v.split("e")[0].split(".")[1]) == 8 for v in text)   # "%.8e"


def test_random_model_and_species_selection(tmp_path):
    m = synth.random_model(12, 20, 4)
    t = np.linspace(0.5, 2.0, 4)          # t0 != 0: 0 is prepended, then dropped
    cs = [3, 0, 7]
    sol = _run(_write(tmp_path, m, t, cs=cs), tmp_path / "out")
    assert sol.shape == (4, 4) and sol[0, 0] == 0.5
    ref = lsoda.solve(m, t, cs=np.array(cs))
    np.testing.assert_allclose(sol[:, 1:], ref, rtol=2e-5, atol=1e-10)


def test_stoichiometry_three_contributes_no_factor(tmp_path):
    """3A -> B: the generated code has no factor for A, so dA/dt = -3k and
    dB/dt = k exactly (GenerateODEs.py:231-238)."""
    k, a0 = 0.05, 2.0
    m = cases.from_reactions(2, [({0: 3}, {1: 1})], [k], [a0, 0.0])
    t = np.linspace(0.0, 4.0, 5)
    sol = _run(_write(tmp_path, m, t), tmp_path / "out")
    np.testing.assert_allclose(sol[:, 1], a0 - 3 * k * t, rtol=1e-9)
    np.testing.assert_allclose(sol[:, 2], k * t, rtol=1e-9, atol=1e-14)


def test_feed_species_is_frozen_at_its_feed_value(tmp_path):
    m = cases.with_feed()
    t = np.linspace(0.0, 2.0, 5)
    sol = _run(_write(tmp_path, m, t), tmp_path / "out")
    assert np.all(sol[:, 2] == 0.5)
    ref = lsoda.solve(m, t)
    np.testing.assert_allclose(sol[:, 1:], ref, rtol=2e-5, atol=1e-10)


def test_fitness_mode(tmp_path, capsys):
    m = cases.chain()
    t = np.linspace(0.0, 2.0, 5)
    d = _write(tmp_path, m, t)
    sim = _run(d, tmp_path / "out")
    target = sim.copy()
    target[:, 1:] *= 1.1
    np.savetxt(d / "ts_matrix", target)
    capsys.readouterr()
    assert lsodasim.main([str(d), str(tmp_path / "out2"), "-f"]) == 0
    acc = float(capsys.readouterr().out.strip().splitlines()[-1])
    nz = np.count_nonzero(target[:, 1:])
    assert acc == pytest.approx(nz * (0.1 / 1.1), rel=1e-6)
    assert not (tmp_path / "out2" / "ODESol").exists()


def test_stochastic_conversion_constants():
    L = np.array([[1, 0, 0], [1, 1, 0], [2, 0, 0]])
    k = lsodasim.convert_constants([1.0, 2.0, 3.0], L, 10.0)
    assert list(k) == [1.0, 20.0, 15.0]
    with pytest.raises(ValueError):
        lsodasim.convert_constants([1.0], np.array([[1, 1, 1]]), 10.0)


@pytest.mark.parametrize("seed", [0, 1])
def test_jacobian_matches_central_differences(seed):
    m = synth.random_model(10, 16, seed)
    sysm = lsodasim.FastGeneratedSystem(m.L.astype(int), m.Rm.astype(int), m.k, m.feed)
    slow = lsodasim.GeneratedSystem(m.L.astype(int), m.Rm.astype(int), m.k, m.feed)
    y = np.random.default_rng(seed).uniform(0.3, 1.5, m.S)
    J = sysm.jac(y)
    np.testing.assert_allclose(J, slow.jac(y), rtol=1e-13, atol=1e-15)
    h = 1e-6
    fd = np.column_stack([(sysm.f(y + h * e) - sysm.f(y - h * e)) / (2 * h) for e in np.eye(m.S)])
    np.testing.assert_allclose(J, fd, rtol=1e-6, atol=1e-8)
