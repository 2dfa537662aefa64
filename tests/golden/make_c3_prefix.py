"""Writes tests/golden/c3_prefix_4096.npz: the oracle's state and counters
after the first two loop iterations (two BDF1 steps) of the 4096 x 4096
stiff workload (synth.stiff(4096, 4096, 0), k logU[1e-6, 1e6], t in [0, 100]
with 200 samples, rtol 1e-6, BDF step regrowth, 8 RKF retries,
newton_iter 100), in the plain FP64 form and in the precision form (FP32 LU +
FP64 refinement, tol 1e-14, 4 sweeps).  The oracle takes about 90 s on one
core for the two forms, so the GPU test compares against these vectors.

    python tests/golden/make_c3_prefix.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from lassie_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402

KW = dict(max_iters=2, newton_iter=100, rtol=1e-6, bdf_regrow=1, max_rkf_retries=8, refine_tol=1e-14,
          refine_max_iter=4)
KEYS = ("loop_iters", "rkf_attempts", "rkf_accepts", "bdf_steps", "newton_iters", "lu_failures", "nan_repairs",
        "samples", "refine_iters", "refine_fallbacks", "t_final", "newdt_final")


def main():
    m = synth.stiff(4096, 4096, 0)
    t = synth.t_vector(0, 100, 200)
    data = {}
    for name, refine in (("fp64", 0), ("fp32_refine", 1)):
        out, st, rc, _, yf = O.solve(m, t, refine=refine, **KW)
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
