"""Writes tests/golden/c3_prefix_4096.npz: the oracle's state and counters
after the first two loop iterations (two BDF1 steps) of the 4096 x 4096
stiff workload (synth.stiff(4096, 4096, 0), k logU[1e-6, 1e6], t in [0, 100]
with 200 samples, rtol 1e-6, BDF step regrowth, 8 RKF retries,
newton_iter 100), in the plain FP64 form, in the precision form (FP32 LU +
FP64 refinement, tol 1e-14, 4 sweeps) and in the unrefined precision form
(refine_max_iter 0: one FP32 solve per Newton iteration).  The oracle takes
about two minutes on one core for the three forms, so the GPU test compares
against these vectors.

    python tests/golden/make_c3_prefix.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from lassie_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402

KW = dict(max_iters=2, newton_iter=100, rtol=1e-6, bdf_regrow=1, max_rkf_retries=8, refine_tol=1e-14)
KEYS = ("loop_iters", "rkf_attempts", "rkf_accepts", "bdf_steps", "newton_iters", "lu_failures", "nan_repairs",
        "samples", "refine_iters", "refine_fallbacks", "t_final", "newdt_final")


def main():
    m = synth.stiff(4096, 4096, 0)
    t = synth.t_vector(0, 100, 200)
    data = {}
    for name, refine, sweeps in (("fp64", 0, 4), ("fp32_refine", 1, 4), ("fp32_unrefined", 1, 0)):
        out, st, rc, _, yf = O.solve(m, t, refine=refine, refine_max_iter=sweeps, **KW)
        n = int(st["samples"])
        data[f"{name}_y"] = yf
        data[f"{name}_out"] = out[:n]
        data[f"{name}_stats"] = np.array([float(st[k]) for k in KEYS])
        data[f"{name}_rc"] = np.array(rc)
        print(name, {k: st[k] for k in KEYS}, flush=True)
    data["keys"] = np.array(KEYS)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "c3_prefix_4096.npz"), **data)


if __name__ == "__main__":
    main()
