"""The LASSIE command line: argument and input-file errors end with the
reference's exit codes before any GPU work (simulator.cu:61-207,
Simulation.cuh:373-391, 1006-1011); on a GPU box the Solution file equals the
oracle's trajectory printed the reference's way ("%.4e\\t")."""
import os
import subprocess

import numpy as np
import pytest

from lassie_amd import synth
from oracle import oracle as O
from tests import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "lassie_amd", "LASSIE")


def run(*args):
    p = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=120)
    rc = p.returncode
    return (rc - 256 if rc > 127 else rc), p.stdout + p.stderr


def model_folder(tmp_path, m=None, name="model", t=None, extra=None, cs=None):
    m = m or synth.small_nonstiff(0)
    d = tmp_path / name
    synth.write_folder(m, str(d), synth.t_vector(0, 2, 5) if t is None else t, cs=cs, extra=extra)
    return d


def test_argument_errors():
    assert run()[0] == -3
    assert run("x")[0] == -2


def test_missing_input_folder(tmp_path):
    rc, out = run(str(tmp_path / "nope"), str(tmp_path / "out"))
    assert rc == -1 and "input folder does not exist" in out


def test_missing_mandatory_files(tmp_path):
    d = model_folder(tmp_path)
    os.remove(d / "c_vector")
    os.remove(d / "M_0")
    rc, out = run(str(d), str(tmp_path / "out"))
    assert rc == -4
    assert "kinetic constants vector" in out and "initial conditions" in out


def test_bad_output_and_precision_flags(tmp_path):
    d = model_folder(tmp_path)
    assert run(str(d), "-v1")[0] == -2
    assert run(str(d), str(tmp_path / "o"), "-p")[0] == -5
    assert run(str(d), str(tmp_path / "o"), "-p", "half")[0] == -6


def test_modelkind_and_volume_errors(tmp_path):
    d = model_folder(tmp_path, extra={"modelkind": "sideways"})
    assert run(str(d), str(tmp_path / "o"))[0] == -7
    d2 = model_folder(tmp_path, name="m2", extra={"modelkind": "stochastic"})
    assert run(str(d2), str(tmp_path / "o2"))[0] == -8


def test_stochastic_rejects_three_reactants(tmp_path):
    m = cases.from_reactions(3, [({0: 1, 1: 1, 2: 1}, {})], [1.0], [1.0, 1.0, 1.0])
    d = model_folder(tmp_path, m, extra={"modelkind": "stochastic", "volume": "1e-15"})
    assert run(str(d), str(tmp_path / "o"))[0] == -12


def test_output_folder_and_sampling_file_are_created(tmp_path):
    d = model_folder(tmp_path, cs=[3, 1])
    out = tmp_path / "out"
    run(str(d), str(out))
    assert (out / "output").is_dir()
    assert (out / "output" / "SpeciesSampling").read_text() == "Saving species' dynamics: \n3; 1; \n"


def _fmt(t, out):
    return "".join("%.4e\t" % ti + "".join("%.4e\t" % v for v in row) + "\n" for ti, row in zip(t, out))


@pytest.mark.gpu
def test_solution_file_matches_oracle(tmp_path, gpu_lib):
    m = synth.small_nonstiff(0)
    t = synth.t_vector(0, 10, 100)
    d = model_folder(tmp_path, m, t=t, cs=[0, 5, 31])
    print(f"[cli] 32x48 model, faithful defaults", flush=True)
    rc, log = run(str(d), str(tmp_path / "out"))
    assert rc == 0, log
    o_out, o_st, o_rc, prog, _ = O.solve(m, t, cs=np.array([0, 5, 31]), progress_cap=1 << 20)
    assert (tmp_path / "out" / "output" / "Solution").read_text() == _fmt(t, o_out)
    lines = (tmp_path / "out" / "output" / "ProgressSimulation").read_text().splitlines()
    assert lines == ["%.4e" % v for v in prog]


@pytest.mark.gpu
def test_precision_float_runs_the_single_precision_loop(tmp_path, gpu_lib):
    """-p float: Simulation<float> (simulator.cu:259-304); the file holds the
    single-precision oracle's trajectory, times as float values."""
    m = synth.small_nonstiff(0)
    t = synth.t_vector(0, 10, 100)
    d = model_folder(tmp_path, m, t=t, cs=[0, 5, 31])
    print("[cli] 32x48 model, -p float", flush=True)
    rc, log = run(str(d), str(tmp_path / "out"), "-p", "float")
    assert rc == 0, log
    o_out, o_st, o_rc, prog, _ = O.solve32(m, t, cs=np.array([0, 5, 31]), progress_cap=1 << 20)
    t32 = t.astype(np.float32).astype(np.float64)
    assert (tmp_path / "out" / "output" / "Solution").read_text() == _fmt(t32, o_out)
    lines = (tmp_path / "out" / "output" / "ProgressSimulation").read_text().splitlines()
    assert lines == ["%.4e" % v for v in prog]


@pytest.mark.gpu
def test_sbml_import_then_lassie(tmp_path, gpu_lib):
    """SBML -> BioSimWare (lassie_amd/sbml.py) -> LASSIE: the Solution file
    equals the oracle's trajectory of the converted model (its constant
    species is a feed).  Kept with the other command-line runs: a child
    process cannot make a cooperative launch while this process holds the
    GPU's cooperative queue (after its own solves)."""
    from lassie_amd import sbml
    from tests.test_sbml import L2_DOC
    m = sbml.parse(L2_DOC)
    t = np.linspace(0.0, 2.0, 11)
    sbml.write_biosimware(m, str(tmp_path / "in"), t=t)
    print("[cli] SBML toy model", flush=True)
    rc, log = run(str(tmp_path / "in"), str(tmp_path / "out"))
    assert rc == 0, log
    model = sbml.to_model(m)
    model.k = np.loadtxt(tmp_path / "in" / "c_vector")          # as the reader sees them ("%e")
    model.x0 = np.loadtxt(tmp_path / "in" / "M_0", delimiter="\t")
    out, st, o_rc, _, _ = O.solve(model, t, max_iters=2_000_000)
    assert o_rc == 0
    assert (tmp_path / "out" / "output" / "Solution").read_text() == _fmt(t, out)


@pytest.mark.gpu
def test_stochastic_conversion_matches_oracle(tmp_path, gpu_lib):
    m = cases.mixed()
    vol, avo = 1e-15, 6.022e23
    m_num = synth.model_from_dense(m.L, m.Rm, m.k, m.x0 * vol * avo)
    t = synth.t_vector(0, 3, 7)
    d = model_folder(tmp_path, m_num, t=t, extra={"modelkind": "stochastic", "volume": "%.17g" % vol})
    rc, log = run(str(d), str(tmp_path / "out"))
    assert rc == 0, log
    # Simulation.cuh:983-1031: x0 / (V N_A); bimolecular k * V N_A; homodimer k * V N_A / 2
    conv = vol * avo
    k = m.k.copy()
    for r in range(m.R):
        distinct = int(np.sum(m.L[r] > 0))
        if np.any(m.L[r] >= 2):
            k[r] = (k[r] * conv) / 2.0
        elif distinct > 1:
            k[r] = k[r] * conv
    m_conc = synth.model_from_dense(m.L, m.Rm, k, (m.x0 * vol * avo) / conv)
    o_out, _, o_rc, _, _ = O.solve(m_conc, t)
    assert (tmp_path / "out" / "output" / "Solution").read_text() == _fmt(t, o_out * (vol * avo))


def test_verbose_prints_odes_and_jacobian(tmp_path):
    """-v3 prints the symbolic ODEs and Jacobian before any GPU work
    (Simulation.cuh:433-449).  Model: 2A -> B (k0), A + B -> C (k1), 0 -> A (k2);
    expected text written out by hand from printOdes / printJac_ (a species of
    order 2 contributes "(c*2)*k[r]*x[g]" with no explicit "+")."""
    m = cases.from_reactions(3, [({0: 2}, {1: 1}), ({0: 1, 1: 1}, {2: 1}), ({}, {0: 1})], [1.0, 2.0, 3.0],
                             [1.0, 0.5, 0.0])
    d = model_folder(tmp_path, m)
    rc, out = run(str(d), str(tmp_path / "out"), "-v3")
    odes = ["dx[0] = -2*k[0]*x[0]*x[0]-1*k[1]*x[0]*x[1]+1*k[2]",
            "dx[1] = +1*k[0]*x[0]*x[0]-1*k[1]*x[0]*x[1]",
            "dx[2] = +1*k[1]*x[0]*x[1]"]
    jac = ["dx[0][0] = -4*k[0]*x[0]-1*k[1]*x[1]", "dx[0][1] = -1*k[1]*x[0]", "dx[0][2] = 0",
           "dx[1][0] = 2*k[0]*x[0]-1*k[1]*x[1]", "dx[1][1] = -1*k[1]*x[0]", "dx[1][2] = 0",
           "dx[2][0] = +1*k[1]*x[1]", "dx[2][1] = +1*k[1]*x[0]", "dx[2][2] = 0"]
    lines = out.splitlines()
    i = lines.index("ODEs:")
    assert lines[i + 1:i + 4] == odes
    j = lines.index("Jacobian:")
    assert lines[j + 1:j + 10] == jac


def test_fed_species_print_no_ode_and_zero_jacobian_columns(tmp_path):
    m = cases.with_feed()
    d = model_folder(tmp_path, m)
    rc, out = run(str(d), str(tmp_path / "out"), "-v3")
    lines = out.splitlines()
    i = lines.index("ODEs:")
    assert not any(l.startswith("dx[1] =") for l in lines[i + 1:i + 4])  # species 1 is fed
    assert all(f"dx[{r}][1] = 0" in lines for r in range(3))


def block_thread_text(sm, thread, kind, save):
    """infoBlockThread section of the reference's launch rule
    (returnBlockThread, Simulation.cuh:1225-1277): threads per block =
    min(256, ceil(thread / sm)) rounded up to a multiple of 32, in float."""
    s = (f"{kind} \n" if save == 0 else f"\n\n{kind} \n")
    s += f"Allocating {thread} threads over {sm} available streaming multiprocessors\n"
    put = min(256, int(np.ceil(np.float32(thread) / np.float32(sm))))
    s += f"Putative TPB: {put} \n"
    tpb = put
    if put % 32:
        s += "Adapting to warp size\n"
        tpb = (put // 32 + 1) * 32
    blocks = int(np.ceil(np.float32(thread) / np.float32(tpb)))
    return s + f"Final TPB: {tpb} \nNumber of used blocks: {blocks} \nTotal threads spawned: {tpb * blocks} \n"


def test_block_thread_rule_known_answers():
    """Hand-worked cases of the rule on 256 compute units: 32 species (ODE),
    their 1024 Jacobian entries, 3 sampled species, and 4096^2 entries."""
    assert block_thread_text(256, 32, "ODE", 0) == (
        "ODE \nAllocating 32 threads over 256 available streaming multiprocessors\nPutative TPB: 1 \n"
        "Adapting to warp size\nFinal TPB: 32 \nNumber of used blocks: 1 \nTotal threads spawned: 32 \n")
    assert block_thread_text(256, 1024, "JAC", 1) == (
        "\n\nJAC \nAllocating 1024 threads over 256 available streaming multiprocessors\nPutative TPB: 4 \n"
        "Adapting to warp size\nFinal TPB: 32 \nNumber of used blocks: 32 \nTotal threads spawned: 1024 \n")
    assert block_thread_text(256, 4096 * 4096, "JAC", 1).endswith(
        "Putative TPB: 256 \nFinal TPB: 256 \nNumber of used blocks: 65536 \nTotal threads spawned: 16777216 \n")
    assert "Adapting" not in block_thread_text(256, 8192, "ODE", 0)  # 32 per block exactly


@pytest.mark.gpu
def test_verbose_two_reports_launch_and_progress(tmp_path, gpu_lib):
    """-v2: the THREAD INFORMATIONS block, the model summary, the start and
    end banners and the quarter marks as samples are stored, in the
    reference's order (Simulation.cuh:602-620, 655-657, 679-687, 934-937,
    973-977); infoBlockThread holds the ODE, Jacobian and sampling records."""
    m = synth.small_nonstiff(0)
    t = synth.t_vector(0, 2, 8)
    d = model_folder(tmp_path, m, t=t, cs=[0, 5, 31])
    print("[cli] 32x48 model, -v2", flush=True)
    rc, out = run(str(d), str(tmp_path / "out"), "-v2")
    assert rc == 0, out
    info = (tmp_path / "out" / "output" / "infoBlockThread").read_text()
    sm = int(info.split("threads over ")[1].split()[0])
    assert info == (block_thread_text(sm, 32, "ODE", 0) + block_thread_text(sm, 32 * 32, "JAC", 1)
                    + block_thread_text(sm, 3, "DINAM-SUBSPEC", 1))
    ode = block_thread_text(sm, 32, "ODE", 0).splitlines()
    jac = block_thread_text(sm, 1024, "JAC", 1).splitlines()
    blocks = lambda lines: int(lines[-2].split(":")[1])
    tpb = lambda lines: int(lines[-3].split(":")[1])
    expect = ["THREAD INFORMATIONS:", f"Number of blocks: {blocks(ode)}", f"Number of threads: {tpb(ode)}",
              f"Number of Jacobian blocks: {blocks(jac)}", f"Number of Jacobian threads: {tpb(jac)}", "",
              "***************************************************************",
              "The model has 32 species and 48 reactions", "",
              "***************************************************************", "modelkind = deterministic",
              "", "", "###############################################################", "Start ode solver", "",
              "****Start simulation***", "****1/4 of simulation***", "****1/2 of simulation***",
              "****3/4 of simulation***", "****End simulation***", "", "End ode solver",
              "###############################################################"]
    lines = out.splitlines()
    i = lines.index("THREAD INFORMATIONS:")
    assert lines[i:i + len(expect)] == expect


@pytest.mark.gpu
def test_progress_file_longer_than_the_first_buffer(tmp_path, gpu_lib):
    """A run with more loop iterations than the CLI's first progress buffer
    (set to 50 lines here) writes every line: the solve is repeated with room
    for all of them, and the file equals the oracle's log."""
    m = synth.small_nonstiff(0)
    t = synth.t_vector(0, 10, 100)
    d = model_folder(tmp_path, m, t=t, cs=[0, 5, 31])
    print("[cli] 32x48 model, 50-line first progress buffer", flush=True)
    p = subprocess.run([EXE, str(d), str(tmp_path / "out")], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, LASSIE_PROGRESS_BUFFER="50"))
    assert p.returncode == 0, p.stdout + p.stderr
    o_out, o_st, _, prog, _ = O.solve(m, t, cs=np.array([0, 5, 31]), progress_cap=1 << 20)
    assert o_st["loop_iters"] > 50
    lines = (tmp_path / "out" / "output" / "ProgressSimulation").read_text().splitlines()
    assert lines == ["%.4e" % v for v in prog]
    assert (tmp_path / "out" / "output" / "Solution").read_text() == _fmt(t, o_out)
