"""Single-precision loop (`-p float`, LASSIE_PREC_FP32_LOOP) on the GPU vs the
single-precision oracle (oracle/lassie_oracle_f32.c).

Bit-exact, as for the FP64 loop: float state, time and LU, with the double
accumulators and double-literal expressions of the reference's
Simulation<float>.  Covers the LDS-resident and global control paths, every
RHS/Jacobian branch (hand models), feed species, a singular Newton matrix,
ragged sizes through the panel LU (16-column panels with vector strip
updates), and the 1024-species stiff workload over a prefix.
"""
import numpy as np
import pytest

from lassie_amd import _lib, synth
from oracle import oracle as O
from tests import cases

pytestmark = pytest.mark.gpu

COUNTERS = ("loop_iters", "rkf_attempts", "rkf_accepts", "bdf_steps", "newton_iters",
            "lu_failures", "nan_repairs", "samples")
BENCH = dict(rtol=1e-6, bdf_regrow=1, max_rkf_retries=8)


def run_both(gpu_lib, m, t, cs=None, cap=0, reserved=0, **opts):
    print(f"[fp32] S={m.S} R={m.R} nt={len(t)} opts={opts} cap={cap}", flush=True)
    dm = gpu_lib.DeviceModel(m)
    out_g, yf_g, st_g, rc_g, _ = dm.solve(t, cs=cs, raise_on_maxiter=False, max_iters=cap,
                                          precision=_lib.PREC_FP32_LOOP, reserved=reserved, **opts)
    dm.close()
    out_o, st_o, rc_o, _, yf_o = O.solve32(m, t, cs=cs, max_iters=cap, **opts)
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
    # every value is a float widened to double
    assert np.array_equal(out_g[:n], out_g[:n].astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("name", ["decay", "reversible", "chain", "dimer", "mixed", "with_feed",
                                  "autocatalytic_singular"])
@pytest.mark.parametrize("opts", [dict(), BENCH], ids=["faithful", "bench"])
def test_hand_models_fp32(gpu_lib, name, opts):
    m = getattr(cases, name)()
    g, o = run_both(gpu_lib, m, synth.t_vector(0, 5, 21), cap=200_000, **opts)
    assert_identical(g, o)


def test_singular_newton_matrix_fp32(gpu_lib):
    """A -> 2A with k = 1/h: 1 - hk rounds to 0 in float at the first BDF step."""
    m = cases.autocatalytic_singular(h=0.02)
    g, o = run_both(gpu_lib, m, synth.t_vector(0, 5, 21), cap=200_000, be_step=0.02, **BENCH)
    assert o[2]["lu_failures"] >= 1
    assert_identical(g, o)


def test_small_nonstiff_fp32(gpu_lib):
    m = synth.small_nonstiff(0)
    g, o = run_both(gpu_lib, m, synth.t_vector(0, 10, 100), cap=2_000_000, rtol=1e-4)
    assert o[2]["samples"] == 100
    assert_identical(g, o)


@pytest.mark.parametrize("S,seed", [(16, 1000), (48, 1001), (64, 1002)])
def test_stiff_instances_fp32(gpu_lib, S, seed):
    m = synth.stiff(S, 2 * S, seed, (1e-4, 1e4))
    g, o = run_both(gpu_lib, m, synth.t_vector(0, 50, 100), cap=400_000, **BENCH)
    assert o[2]["bdf_steps"] > 0
    assert_identical(g, o)


@pytest.mark.parametrize("S", [17, 100, 257, 513])
def test_ragged_sizes_prefix_fp32(gpu_lib, S):
    m = synth.stiff(S, S, 7)
    g, o = run_both(gpu_lib, m, synth.t_vector(0, 100, 200), cap=40, **BENCH)
    assert o[2]["bdf_steps"] > 10
    assert_identical(g, o)


def test_global_control_path_fp32(gpu_lib):
    """opts.reserved bit 1 keeps the control group's vectors in global memory."""
    m = synth.stiff(64, 64, 11)
    g, o = run_both(gpu_lib, m, synth.t_vector(0, 20, 50), cap=5000, reserved=2, **BENCH)
    assert_identical(g, o)


def test_stiff_1024_prefix_fp32(gpu_lib):
    m = synth.stiff(1024, 1024, 0)
    g, o = run_both(gpu_lib, m, synth.t_vector(0, 100, 200), cap=10, newton_iter=100, **BENCH)
    assert o[2]["bdf_steps"] >= 4
    assert_identical(g, o)


def test_eight_column_panels_2100_prefix_fp32(gpu_lib):
    """n > 2048: 8-column float panels, 16 rows per thread."""
    m = synth.stiff(2100, 2100, 0)
    g, o = run_both(gpu_lib, m, synth.t_vector(0, 100, 200), cap=4, newton_iter=100, **BENCH)
    assert o[2]["bdf_steps"] >= 2
    assert_identical(g, o)


def test_fp64_unchanged_by_fp32_tables(gpu_lib):
    """A float solve, then a double solve on the same handle: the double solve
    still matches the FP64 oracle (the float coefficient tables are separate)."""
    m = cases.mixed()
    t = synth.t_vector(0, 5, 21)
    dm = gpu_lib.DeviceModel(m)
    dm.solve(t, precision=_lib.PREC_FP32_LOOP, max_iters=200_000)
    out, yf, st, rc, _ = dm.solve(t, max_iters=200_000)
    dm.close()
    o_out, o_st, _, _, o_y = O.solve(m, t, max_iters=200_000)
    np.testing.assert_array_equal(out, o_out)
    np.testing.assert_array_equal(yf, o_y)
