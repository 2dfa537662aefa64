"""Plain FP64 GPU path vs the oracle (CPU restatement of LASSIE's integrator).

The bar is bit-exact: the sampled trajectory, the final state and every
controller counter must equal the oracle's, because the accept/reject, BDF
switch and Newton-stop decisions are threshold tests that any ulp difference
could flip.  Models small enough for the oracle run whole trajectories; the
1024-species stiff workload is compared over a capped prefix of the loop.
"""
import numpy as np
import pytest

from lassie_amd import synth
from oracle import oracle as O
from tests import cases

pytestmark = pytest.mark.gpu

COUNTERS = ("loop_iters", "rkf_attempts", "rkf_accepts", "bdf_steps", "newton_iters",
            "lu_failures", "nan_repairs", "samples")


def run_both(gpu_lib, m, t, cs=None, cap=0, **opts):
    print(f"[parity] S={m.S} R={m.R} nt={len(t)} opts={opts} cap={cap}", flush=True)
    dm = gpu_lib.DeviceModel(m)
    out_g, yf_g, st_g, rc_g, _ = dm.solve(t, cs=cs, raise_on_maxiter=False, max_iters=cap, **opts)
    dm.close()
    out_o, st_o, rc_o, _, yf_o = O.solve(m, t, cs=cs, max_iters=cap, **opts)
    return (out_g, yf_g, st_g, rc_g), (out_o, yf_o, st_o, rc_o)


def assert_identical(g, o):
    out_g, yf_g, st_g, rc_g = g
    out_o, yf_o, st_o, rc_o = o
    assert (rc_g == 0) == (rc_o == 0), (rc_g, rc_o)
    for key in COUNTERS:
        assert st_g[key] == st_o[key], (key, st_g[key], st_o[key])
    assert st_g["t_final"] == st_o["t_final"]
    assert st_g["newdt_final"] == st_o["newdt_final"]
    np.testing.assert_array_equal(yf_g, yf_o)
    n = int(st_o["samples"])
    np.testing.assert_array_equal(out_g[:n], out_o[:n])


FAITHFUL = dict()
BENCH = dict(rtol=1e-6, bdf_regrow=1, max_rkf_retries=8)


@pytest.mark.parametrize("opts", [FAITHFUL, dict(rtol=1e-6)], ids=["faithful", "rtol1e-6"])
def test_small_nonstiff_32x48(gpu_lib, opts):
    m = synth.small_nonstiff(0)
    g, o = run_both(gpu_lib, m, synth.t_vector(0, 10, 100), cap=2_000_000, **opts)
    assert o[2]["samples"] == 100
    assert_identical(g, o)


@pytest.mark.parametrize("name", ["decay", "reversible", "chain", "dimer", "mixed", "with_feed",
                                  "autocatalytic_singular", "hub"])
@pytest.mark.parametrize("opts", [FAITHFUL, BENCH], ids=["faithful", "bench"])
def test_hand_models(gpu_lib, name, opts):
    m = getattr(cases, name)()
    g, o = run_both(gpu_lib, m, synth.t_vector(0, 5, 21), cap=200_000, **opts)
    assert_identical(g, o)
    if name == "autocatalytic_singular":
        assert o[2]["lu_failures"] >= 1


@pytest.mark.parametrize("S,seed", [(16, 1000), (32, 1001), (64, 1002)])
def test_stiff_batch_instances(gpu_lib, S, seed):
    m = synth.stiff(S, 2 * S, seed, (1e-4, 1e4))
    g, o = run_both(gpu_lib, m, synth.t_vector(0, 50, 100), cap=2_000_000, **BENCH)
    assert o[3] == 0 and o[2]["samples"] == 100
    assert_identical(g, o)


@pytest.mark.parametrize("S", [17, 100, 257, 300, 513])
def test_ragged_sizes_prefix(gpu_lib, S):
    """Sizes off the panel grid; stiff, so BDF/LU run from the first steps."""
    m = synth.stiff(S, S, 7)
    g, o = run_both(gpu_lib, m, synth.t_vector(0, 100, 200), cap=40, **BENCH)
    assert o[2]["bdf_steps"] > 10
    assert_identical(g, o)


def test_cs_subset_and_offset_times(gpu_lib):
    m = synth.small_nonstiff(3)
    cs = np.array([5, 0, 31, 5], np.int32)
    t = np.linspace(0.5, 4.0, 8)
    g, o = run_both(gpu_lib, m, t, cs=cs, cap=2_000_000, rtol=1e-6)
    assert_identical(g, o)


def test_maxiter_status(gpu_lib):
    m = synth.stiff(64, 64, 3)
    dm = gpu_lib.DeviceModel(m)
    _, _, st, rc, _ = dm.solve(synth.t_vector(0, 100, 200), raise_on_maxiter=False, max_iters=25)
    assert rc == gpu_lib.LASSIE_ERR_MAXITER and st["loop_iters"] == 25
    with pytest.raises(gpu_lib.LassieError):
        dm.solve(synth.t_vector(0, 100, 200), max_iters=25)


def test_stiff_1024_prefix(gpu_lib):
    """The 1024 x 1024 stiff workload (k logU[1e-6, 1e6]), first loop iterations."""
    m = synth.stiff(1024, 1024, 0)
    g, o = run_both(gpu_lib, m, synth.t_vector(0, 100, 200), cap=10, newton_iter=100, **BENCH)
    assert o[2]["bdf_steps"] >= 4
    assert_identical(g, o)


def test_eight_column_panels_2100_prefix(gpu_lib):
    """n > 2048: the FP64 kernel with 8-column panels and the vector strip
    update (the 4096-species configuration's kernel), first loop iterations."""
    m = synth.stiff(2100, 2100, 0)
    g, o = run_both(gpu_lib, m, synth.t_vector(0, 100, 200), cap=4, newton_iter=100, **BENCH)
    assert o[2]["bdf_steps"] >= 2
    assert_identical(g, o)


def test_overflowing_rate_prefix(gpu_lib):
    """Infinite Newton-matrix entries in the first panels (cases.overflowing_rate,
    S = 300, two rows per thread): the panels that gather their rows meet
    pivot rows with inf and fall back to the whole panel; every BDF step
    repairs NaNs.  Both controllers, first 30 loop iterations."""
    m = cases.overflowing_rate()
    for opts in (BENCH, FAITHFUL):
        g, o = run_both(gpu_lib, m, synth.t_vector(0, 100, 200), cap=30, newton_iter=100, **opts)
        assert o[2]["bdf_steps"] >= 10 and o[2]["nan_repairs"] >= 10
        assert_identical(g, o)

