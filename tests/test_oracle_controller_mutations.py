"""The controller tests of test_oracle_controller.py each catch the mistake
they are written against.  For every rule, a copy of the oracle's source with
that rule changed is compiled (gcc, the oracle's own flags) into a temporary
directory, the oracle module is pointed at it, and the rule's test must fail;
the unchanged oracle passes every one of them (test_oracle_controller.py).
The mistakes are planted in the oracle on the CPU only."""
import os
import shutil
import subprocess

import pytest

from oracle import oracle as O
from tests import test_oracle_controller as T

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")

# (rule, text in lassie_oracle.c, replacement, the test that must fail)
MUTATIONS = [
    ("time advances by the step taken (t += dt_old)",
     "            if ((t + d_temp) > t1) dt = fabs(t1 - t);\n            else dt = d_temp;\n            t = t + dt;",
     "            const double dt_taken = dt;\n            if ((t + d_temp) > t1) dt = fabs(t1 - t);\n"
     "            else dt = d_temp;\n            t = t + dt_taken;",
     T.test_accept_advances_time_by_the_next_step),
    ("no clamp of the next RKF45 step to t1",
     "            if ((t + d_temp) > t1) dt = fabs(t1 - t);",
     "            if (0 && (t + d_temp) > t1) dt = fabs(t1 - t);",
     T.test_accept_clamps_the_next_step_to_the_sample_time),
    ("sampling test relative to t1",
     "        if (fabs(t1 - t) < 1e-5) {",
     "        if (fabs(t1 - t) < 1e-5 * fabs(t1)) {",
     T.test_sampling_tolerance_is_absolute_1e_minus_5),
    ("stiffness switch relative to dt_old",
     "fabs(dt_old - dt) < 1e-3",
     "fabs(dt_old - dt) < 1e-3 * dt_old",
     T.test_stiffness_switch_is_absolute_1e_minus_3),
    ("no clamp of the BDF1 step to t1",
     "    if ((t + dt) > t1) dt = fabs(t1 - t);\n    memcpy(w->yv, y1",
     "    if (0 && (t + dt) > t1) dt = fabs(t1 - t);\n    memcpy(w->yv, y1",
     T.test_bdf_step_is_cut_to_the_sample_time),
    ("Newton norm relative to the state",
     "        errN = sqrt(orc_normsq(S, w->x));",
     "        errN = sqrt(orc_normsq(S, w->x)) / fabs(w->yv[0]);",
     T.test_newton_stops_on_the_absolute_norm),
]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc to build the oracle copies")
@pytest.mark.parametrize("rule,old,new,test", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_planted_mistake_is_caught(tmp_path, rule, old, new, test):
    for f in ("lassie_oracle.c", "lassie_oracle_adjoint.c", "lassie_oracle_f32.c", "lassie_oracle.h"):
        shutil.copy(os.path.join(ORACLE, f), tmp_path / f)
    src = (tmp_path / "lassie_oracle.c").read_text()
    assert src.count(old) == 1, rule
    (tmp_path / "lassie_oracle.c").write_text(src.replace(old, new))
    so = tmp_path / "liblassie_oracle_mutant.so"
    subprocess.run(["gcc", "-O3", "-fPIC", "-std=c11", "-ffp-contract=off", "-mfma", "-fno-fast-math", "-shared",
                    "-o", str(so), "lassie_oracle.c", "lassie_oracle_adjoint.c", "lassie_oracle_f32.c", "-lm"],
                   cwd=tmp_path, check=True)
    O.use_library(str(so))
    try:
        with pytest.raises(AssertionError):
            test()
    finally:
        O.use_library(None)
