"""The seeded generator of the synthetic workloads: structure, ranges and
determinism (the same seed gives the same model on every host)."""
import numpy as np

from lassie_amd import synth


def test_generator_structure():
    m = synth.random_model(64, 200, 5, (1e-3, 1e1))
    nreact = (m.L > 0).sum(axis=1)
    nprod = (m.Rm > 0).sum(axis=1)
    assert set(np.unique(nreact)) <= {1, 2} and set(np.unique(nprod)) <= {0, 1, 2}
    assert m.L.max() == 1 and m.Rm.max() == 1          # stoichiometry 1, order <= 2
    assert np.all(m.k >= 1e-3) and np.all(m.k <= 1e1)
    assert np.all((m.x0 >= 0) & (m.x0 <= 1))


def test_generator_is_deterministic():
    a, b = synth.stiff(128, 128, 3), synth.stiff(128, 128, 3)
    for f in ("L", "Rm", "k", "x0"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f))
    assert not np.array_equal(a.k, synth.stiff(128, 128, 4).k)
    assert np.all(a.x0 >= 1e-6) and np.all(a.x0 <= 1.0)
    assert a.k.min() >= 1e-6 and a.k.max() <= 1e6


def test_uneven_batch_sizes():
    ms = synth.uneven_batch(60, 1000)
    sizes = {m.S for m in ms}
    assert sizes <= {16, 32, 64, 128, 256, 512} and len(sizes) >= 4
    assert all(m.R == 2 * m.S for m in ms)


def test_folder_roundtrip(tmp_path):
    m = synth.small_nonstiff(2)
    synth.write_folder(m, str(tmp_path / "f"), synth.t_vector(0, 1, 3))
    L = np.loadtxt(tmp_path / "f" / "left_side", delimiter="\t", dtype=int)
    np.testing.assert_array_equal(L, m.L)
    np.testing.assert_array_equal(np.loadtxt(tmp_path / "f" / "c_vector"), m.k)
