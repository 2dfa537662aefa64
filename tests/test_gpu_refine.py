"""Single-model precision form on the GPU (LASSIE_PREC_FP32_REFINE): the FP64
integrator with the Newton matrix rounded to FP32 and factored on the grid,
FP64 iterative refinement of every Newton solve, and the FP64 LU of the
step's matrix when refinement misses its target.

The bar is the plain path's: bit-exact against the oracle's precision form
(orc_refine_solve inside orc_solve), counters included, because the
refinement sweeps are specified operation by operation (FP32 solve, x + d,
fma-chain residual in column order, componentwise test).  Sizes cover the
16-column panels (n <= 2048) and the 8-column panels of the 4096-species
configuration; large models run a capped prefix of the loop.
"""
import numpy as np
import pytest

from lassie_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

COUNTERS = ("loop_iters", "rkf_attempts", "rkf_accepts", "bdf_steps", "newton_iters",
            "lu_failures", "nan_repairs", "samples", "refine_iters", "refine_fallbacks")
BENCH = dict(rtol=1e-6, bdf_regrow=1, max_rkf_retries=8)
REFINE = 1


def run_both(gpu_lib, m, t, cap=0, refine_tol=1e-14, refine_max_iter=4, **opts):
    print(f"[refine] S={m.S} R={m.R} nt={len(t)} cap={cap} tol={refine_tol} "
          f"sweeps={refine_max_iter} opts={opts}", flush=True)
    dm = gpu_lib.DeviceModel(m)
    out_g, yf_g, st_g, rc_g, _ = dm.solve(t, raise_on_maxiter=False, max_iters=cap,
                                          precision=gpu_lib.PREC_FP32_REFINE, refine_tol=refine_tol,
                                          refine_max_iter=refine_max_iter, **opts)
    dm.close()
    out_o, st_o, rc_o, _, yf_o = O.solve(m, t, max_iters=cap, refine=REFINE, refine_tol=refine_tol,
                                         refine_max_iter=refine_max_iter, **opts)
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


@pytest.mark.parametrize("S,seed,cap", [(17, 21, 0), (48, 3, 0), (100, 7, 60), (300, 7, 40),
                                        (513, 7, 30)])
def test_mild_steps_refine_without_fallback(gpu_lib, S, seed, cap):
    """BDF steps of 1e-3 on stiff models: refinement converges in every
    Newton solve; whole trajectories for the small sizes, capped prefixes
    for the others (ragged sizes off the 16-column panel grid)."""
    m = synth.stiff(S, 2 * S, seed, (1e-4, 1e2))
    t = synth.t_vector(0, 1, 11)
    g, o = run_both(gpu_lib, m, t, cap=cap, be_step=1e-3, **BENCH)
    assert o[2]["refine_iters"] > 0 and o[2]["bdf_steps"] > 0
    assert_identical(g, o)


@pytest.mark.parametrize("S,seed,fallbacks", [(64, 1002, 0), (257, 7, 2)])
def test_stiff_steps(gpu_lib, S, seed, fallbacks):
    """k logU[1e-4, 1e4] with the default BDF step; on the 257-species model
    two Newton matrices are too ill-conditioned for FP32 refinement to reach
    1e-14, so those steps fall back to the FP64 LU part-way through the
    Newton loop."""
    m = synth.stiff(S, 2 * S, seed, (1e-4, 1e4))
    g, o = run_both(gpu_lib, m, synth.t_vector(0, 50, 100), cap=300, **BENCH)
    assert o[2]["refine_fallbacks"] == fallbacks
    assert_identical(g, o)


def test_every_step_falls_back(gpu_lib):
    """tol 0 and one sweep: each BDF step runs the FP32 LU, one sweep, then
    the FP64 LU, and the trajectory equals the FP64 form's."""
    m = synth.stiff(100, 200, 5, (1e-4, 1e4))
    t = synth.t_vector(0, 2, 5)
    g, o = run_both(gpu_lib, m, t, refine_tol=0.0, refine_max_iter=1, **BENCH)
    assert_identical(g, o)
    st = o[2]
    assert st["refine_fallbacks"] == st["bdf_steps"] - st["lu_failures"] > 0
    out64, _, _, _, _ = O.solve(m, t, **BENCH)
    np.testing.assert_array_equal(g[0], out64)


def test_singular_newton_matrix(gpu_lib):
    """A -> 2A with k = 1/h: 1 - hk is 0, so the FP32 LU meets a zero pivot,
    the step switches to the FP64 LU, which meets it too, and the step is an
    LU failure (state repaired, BDF step shrunk)."""
    from tests import cases
    m = cases.autocatalytic_singular(h=0.01)
    g, o = run_both(gpu_lib, m, synth.t_vector(0, 5, 21), cap=200_000, be_step=0.01, **BENCH)
    assert o[2]["lu_failures"] >= 1 and o[2]["refine_fallbacks"] >= 1
    assert_identical(g, o)


def test_stiff_1024_prefix(gpu_lib):
    """The 1024 x 1024 stiff workload (k logU[1e-6, 1e6]), first loop iterations."""
    m = synth.stiff(1024, 1024, 0)
    g, o = run_both(gpu_lib, m, synth.t_vector(0, 100, 200), cap=10, newton_iter=100, **BENCH)
    assert o[2]["bdf_steps"] >= 4
    assert_identical(g, o)


def test_eight_column_panels_2100(gpu_lib):
    """n > 2048: 8-column panels, 16 rows per thread in the FP32 panel."""
    m = synth.stiff(2100, 2100, 0)
    g, o = run_both(gpu_lib, m, synth.t_vector(0, 100, 200), cap=4, newton_iter=100, **BENCH)
    assert o[2]["bdf_steps"] >= 2
    assert_identical(g, o)


def test_option_checks(gpu_lib):
    m = synth.stiff(32, 64, 1)
    dm = gpu_lib.DeviceModel(m)
    t = synth.t_vector(0, 1, 3)
    with pytest.raises(gpu_lib.LassieError):
        dm.solve(t, precision=7)
    with pytest.raises(gpu_lib.LassieError):
        dm.solve(t, precision=gpu_lib.PREC_FP32_REFINE, refine_tol=-1.0)
    out, _, st, rc, _ = dm.solve(t, precision=gpu_lib.PREC_FP32_REFINE, **BENCH)
    assert rc == 0 and np.isfinite(out).all()
    dm.close()


def test_c3_4096_prefix_matches_oracle_vectors(gpu_lib):
    """The 4096 x 4096 workload (synth.stiff(4096, 4096, 0)), first two loop
    iterations (two BDF1 steps, each with its 4096^2 Jacobian, LU and Newton
    solves), in the plain FP64 form, in the precision form (FP32 LU, FP64
    refinement, two refinement fallbacks) and in the unrefined precision form
    (refine_max_iter 0): state, samples and counters equal the oracle's bit
    for bit.  The oracle's vectors take about two minutes of CPU, so they come
    from tests/golden/c3_prefix_4096.npz (make_c3_prefix.py)."""
    import os
    gold = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c3_prefix_4096.npz"))
    keys = [str(k) for k in gold["keys"]]
    m = synth.stiff(4096, 4096, 0)
    t = synth.t_vector(0, 100, 200)
    dm = gpu_lib.DeviceModel(m)
    for name, prec, sweeps in (("fp64", gpu_lib.PREC_FP64, 4), ("fp32_refine", gpu_lib.PREC_FP32_REFINE, 4),
                               ("fp32_unrefined", gpu_lib.PREC_FP32_REFINE, 0)):
        print(f"[refine] 4096x4096 prefix, {name}", flush=True)
        out, yf, st, rc, _ = dm.solve(t, raise_on_maxiter=False, max_iters=2, newton_iter=100, precision=prec,
                                      refine_tol=1e-14, refine_max_iter=sweeps, **BENCH)
        assert (rc == 0) == (int(gold[f"{name}_rc"]) == 0)
        for k, v in zip(keys, gold[f"{name}_stats"]):
            assert st[k] == v, (name, k, st[k], v)
        np.testing.assert_array_equal(yf, gold[f"{name}_y"])
        n = int(st["samples"])
        np.testing.assert_array_equal(out[:n], gold[f"{name}_out"])
    dm.close()


# ---- BF16 form (LASSIE_PREC_BF16_REFINE): FP32 factors whose trailing updates
# run on the BF16 matrix cores (operands rounded to BF16, FP32 accumulation).
# The oracle has no BF16 restatement, so these tests pin what the form must
# keep: the FP64 form's results wherever refinement falls back, and the FP64
# form's decisions and accuracy where the Newton matrices are well
# conditioned.

def _solve(gpu_lib, m, t, **kw):
    dm = gpu_lib.DeviceModel(m)
    out, yf, st, rc, _ = dm.solve(t, raise_on_maxiter=False, **kw)
    dm.close()
    return out, yf, st, rc


@pytest.mark.parametrize("S", [100, 300])
def test_bf16_every_step_falls_back(gpu_lib, S):
    """tol 0 and one sweep: each BDF step runs the BF16-updated FP32 LU, one
    sweep, then the FP64 LU, so the trajectory equals the FP64 oracle's."""
    m = synth.stiff(S, 2 * S, 5, (1e-4, 1e4))
    t = synth.t_vector(0, 2, 5)
    print(f"[bf16] fallback S={S}", flush=True)
    out, yf, st, rc = _solve(gpu_lib, m, t, max_iters=400, precision=gpu_lib.PREC_BF16_REFINE, refine_tol=0.0,
                             refine_max_iter=1, **BENCH)
    out64, st64, rc64, _, y64 = O.solve(m, t, max_iters=400, **BENCH)
    assert st["refine_fallbacks"] == st["bdf_steps"] - st["lu_failures"] > 0
    for key in ("loop_iters", "bdf_steps", "newton_iters", "lu_failures", "nan_repairs", "samples"):
        assert st[key] == st64[key], key
    n = int(st64["samples"])
    np.testing.assert_array_equal(out[:n], out64[:n])
    np.testing.assert_array_equal(yf, y64)


@pytest.mark.parametrize("S,seed,cap", [(48, 3, 0), (300, 7, 40), (513, 7, 30)])
def test_bf16_mild_steps_close_to_fp64(gpu_lib, S, seed, cap):
    """BDF steps of 1e-3 on well-conditioned matrices: refinement of the
    BF16-updated factors meets 1e-14 (floor: the species' atol) without
    falling back, the controller takes the FP64 form's decisions, and the
    samples stay within 1e-4 of the FP64 form's, |a - b| / (|b| + atol)."""
    m = synth.stiff(S, 2 * S, seed, (1e-4, 1e2))
    t = synth.t_vector(0, 1, 11)
    kw = dict(max_iters=cap, be_step=1e-3, **BENCH)
    print(f"[bf16] mild S={S} cap={cap}", flush=True)
    ref, _, st64, _ = _solve(gpu_lib, m, t, **kw)
    out, _, st, _ = _solve(gpu_lib, m, t, precision=gpu_lib.PREC_BF16_REFINE, refine_tol=1e-14,
                           refine_max_iter=30, **kw)
    assert st["refine_iters"] >= st["newton_iters"] > 0 and st["refine_fallbacks"] == 0
    for key in ("loop_iters", "bdf_steps", "newton_iters", "nan_repairs", "samples"):
        assert st[key] == st64[key], key
    n = int(st["samples"])
    rel = float((np.abs(out[:n] - ref[:n]) / (np.abs(ref[:n]) + m.atol[None, :])).max()) if n else 0.0
    print(f"[bf16] S={S} rel={rel:.2e} sweeps={st['refine_iters']} newton={st['newton_iters']}", flush=True)
    assert rel <= 1e-4


def test_bf16_eight_column_panels_2100(gpu_lib):
    """n > 2048: 8-column FP32 panels; the BF16 update pads k to 16.  First
    loop iterations of the stiff 2100^2 model (k logU[1e-6, 1e6]) against the
    FP64 form: the same BDF steps, refinement sweeps run (3 solves fall back
    to the FP64 LU), and the final state within 2e-2 of the FP64 form's
    (8.4e-3 measured: BF16's 8-bit mantissa leaves stiff steps' refined
    corrections further from the FP64 ones than FP32 factors do, and the
    Newton loop stops at its own tolerance)."""
    m = synth.stiff(2100, 2100, 0)
    t = synth.t_vector(0, 100, 200)
    kw = dict(max_iters=4, newton_iter=100, **BENCH)
    print("[bf16] 2100 prefix", flush=True)
    _, y64, st64, _ = _solve(gpu_lib, m, t, **kw)
    _, yb, st, _ = _solve(gpu_lib, m, t, precision=gpu_lib.PREC_BF16_REFINE, refine_tol=1e-14,
                          refine_max_iter=30, **kw)
    assert st["bdf_steps"] == st64["bdf_steps"] >= 2 and st["refine_iters"] > 0
    rel = float((np.abs(yb - y64) / (np.abs(y64) + m.atol)).max())
    print(f"[bf16] 2100 rel={rel:.2e} sweeps={st['refine_iters']} fallbacks={st['refine_fallbacks']}", flush=True)
    assert np.isfinite(yb).all() and rel <= 2e-2



# ---- refine_max_iter = 0: one FP32 (or BF16-updated FP32) solve per Newton
# iteration, unrefined; the modified-Newton iteration's FP64 residual refines.

@pytest.mark.parametrize("S,seed,cap,kr", [(48, 3, 0, (1e-4, 1e4)), (300, 7, 60, (1e-6, 1e6)),
                                           (513, 7, 40, (1e-6, 1e6))])
def test_unrefined_matches_oracle(gpu_lib, S, seed, cap, kr):
    """Bit-exact against the oracle's refine_max_iter = 0 form (one FP32
    solve per Newton iteration, x = 0 + (double)d): trajectory, final state
    and every counter, on stiff models (whole trajectory at 48, capped
    prefixes with two and four rows per thread at 300 and 513)."""
    m = synth.stiff(S, 2 * S, seed, kr)
    g, o = run_both(gpu_lib, m, synth.t_vector(0, 50, 100), cap=cap, refine_max_iter=0, **BENCH)
    assert o[2]["refine_iters"] > 0
    assert o[2]["refine_fallbacks"] > 0 or o[2]["refine_iters"] == o[2]["newton_iters"]
    assert_identical(g, o)


def test_unrefined_eight_column_panels_2100(gpu_lib):
    """The 8-column FP32 panels of n > 2048 in the unrefined form, first
    loop iterations of the stiff 2100^2 model, bit-exact against the oracle."""
    m = synth.stiff(2100, 2100, 0)
    g, o = run_both(gpu_lib, m, synth.t_vector(0, 100, 200), cap=4, refine_max_iter=0, newton_iter=100, **BENCH)
    assert o[2]["bdf_steps"] >= 2 and o[2]["refine_iters"] > 0
    assert_identical(g, o)


@pytest.mark.parametrize("S,seed,cap", [(48, 3, 0), (300, 7, 40), (513, 7, 30)])
def test_bf16_unrefined_is_refinement_with_an_unbounded_target(gpu_lib, S, seed, cap):
    """BF16 form on well-conditioned matrices (BDF steps of 1e-3, k logU[1e-4,
    1e2]): refine_max_iter = 0 equals one refinement sweep whose target every
    finite correction meets (refine_tol = 1e300) bit for bit, counters
    included, and the samples stay within 1e-4 of the FP64 form's.  (Where a
    BF16-updated solve overflows, the sweep's residual is NaN and the refined
    form goes on to the FP64 LU, while the unrefined one takes the correction
    and the loop repairs the state: the two then differ by design.)"""
    m = synth.stiff(S, 2 * S, seed, (1e-4, 1e2))
    t = synth.t_vector(0, 1, 11)
    kw = dict(max_iters=cap, be_step=1e-3, **BENCH)
    print(f"[bf16] unrefined S={S} cap={cap}", flush=True)
    a, ya, sa, _ = _solve(gpu_lib, m, t, precision=gpu_lib.PREC_BF16_REFINE, refine_max_iter=0, **kw)
    b, yb, sb, _ = _solve(gpu_lib, m, t, precision=gpu_lib.PREC_BF16_REFINE, refine_tol=1e300,
                          refine_max_iter=4, **kw)
    for key in COUNTERS:
        assert sa[key] == sb[key], key
    np.testing.assert_array_equal(ya, yb)
    np.testing.assert_array_equal(a, b)
    assert sa["refine_iters"] == sa["newton_iters"] > 0 and sa["refine_fallbacks"] == 0
    ref, _, st64, _ = _solve(gpu_lib, m, t, **kw)
    n = int(min(sa["samples"], st64["samples"]))
    rel = float((np.abs(a[:n] - ref[:n]) / (np.abs(ref[:n]) + m.atol[None, :])).max()) if n else 0.0
    print(f"[bf16] unrefined S={S} rel={rel:.2e} newton={sa['newton_iters']}/{st64['newton_iters']}", flush=True)
    assert rel <= 1e-4
