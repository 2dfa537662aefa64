"""SBML -> BioSimWare conversion (lassie_amd/sbml.py, restating gui/SBML2BSW.py).

Pinned by: a hand-written level-2 document whose expected BioSimWare files are
written out here; the reference's own 4096 x 4096 SBML model
(models/copasi_test_4096x4096.sbml, kept gzipped in tests/golden), whose
structure statistics are known and whose kinetic-law formulas must equal the
converted mass-action rates; and the LASSIE reader on the converted folder.
"""
import gzip
import os
import re

import numpy as np
import pytest

from lassie_amd import sbml
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "copasi_test_4096x4096.sbml.gz")

L2_DOC = """<?xml version="1.0" encoding="UTF-8"?>
<sbml xmlns="http://www.sbml.org/sbml/level2/version4" level="2" version="4">
 <model id="toy">
  <listOfCompartments><compartment id="c" size="1"/></listOfCompartments>
  <listOfSpecies>
   <species id="A" name="Alpha" compartment="c" initialAmount="2"/>
   <species id="B" compartment="c" initialConcentration="0.5"/>
   <species id="C" compartment="c" initialAmount="0" constant="true"/>
   <species id="D" initialAmount="1.5"/>
  </listOfSpecies>
  <listOfParameters>
   <parameter id="kf" value="3"/><parameter id="kr" value="0.25"/><parameter id="kd" value="1e-3"/>
  </listOfParameters>
  <listOfReactions>
   <reaction id="bind" reversible="false">
    <listOfReactants><speciesReference species="A"/><speciesReference species="B"/></listOfReactants>
    <listOfProducts><speciesReference species="C" stoichiometry="2"/></listOfProducts>
    <kineticLaw><math xmlns="http://www.w3.org/1998/Math/MathML"><apply><times/><ci>k1</ci><ci>A</ci><ci>B</ci></apply></math>
     <listOfParameters><parameter id="k1" value="1.5"/></listOfParameters></kineticLaw>
   </reaction>
   <reaction id="iso" reversible="true">
    <listOfReactants><speciesReference species="A"/></listOfReactants>
    <listOfProducts><speciesReference species="D"/></listOfProducts>
    <kineticLaw><math xmlns="http://www.w3.org/1998/Math/MathML"><apply><minus/>
     <apply><times/><ci>kf</ci><ci>A</ci></apply><apply><times/><ci>kr</ci><ci>D</ci></apply></apply></math></kineticLaw>
   </reaction>
   <reaction id="decay" reversible="false">
    <listOfReactants><speciesReference species="D" stoichiometry="2"/></listOfReactants>
    <kineticLaw><math xmlns="http://www.w3.org/1998/Math/MathML"><apply><times/><ci>kd</ci><ci>D</ci><ci>D</ci></apply></math></kineticLaw>
   </reaction>
  </listOfReactions>
 </model>
</sbml>
"""


def test_level2_document_to_biosimware_files(tmp_path):
    m = sbml.parse(L2_DOC)
    assert m.names == ["Alpha_in_c", "B_in_c", "C_in_c", "D"]
    assert m.reactions == ["bind", "iso", "iso (reverse)", "decay"]
    sbml.write_biosimware(m, str(tmp_path), t=[0.0, 1.0])
    files = {n: (tmp_path / n).read_text() for n in os.listdir(tmp_path)}
    assert files["M_0"] == "2.000000e+00\t5.000000e-01\t0.000000e+00\t1.500000e+00\n"
    assert files["M_feed"] == "0\t0\t1\t0\n"
    assert files["alphabet"] == "Alpha_in_c\tB_in_c\tC_in_c\tD\t"
    assert files["left_side"] == "1\t1\t0\t0\n1\t0\t0\t0\n0\t0\t0\t1\n0\t0\t0\t2\n"
    assert files["right_side"] == "0\t0\t2\t0\n0\t0\t0\t1\n1\t0\t0\t0\n0\t0\t0\t0\n"
    assert files["c_vector"] == "1.500000e+00\n3.000000e+00\n2.500000e-01\n1.000000e-03\n"
    assert files["boundaries"] == ""
    assert files["t_vector"] == "0\n1\n"


def test_two_global_parameters_imply_a_reverse_and_errors():
    doc = L2_DOC.replace('id="iso" reversible="true"', 'id="iso" reversible="false"')
    m = sbml.parse(doc)
    assert m.reactions[1:3] == ["iso", "iso (reverse)"] and m.warnings
    with pytest.raises(sbml.SBMLError):
        sbml.parse(L2_DOC.replace("<ci>kd</ci>", "<ci>nope</ci>"))
    with pytest.raises(sbml.SBMLError):
        sbml.parse(L2_DOC.replace('species="B"/>', 'species="Z"/>', 1))


@pytest.fixture(scope="module")
def copasi():
    return sbml.parse(FIXTURE)


def test_reference_4096_model_structure(copasi):
    """models/copasi_test_4096x4096.sbml: 4096 species and reactions; 1 or 2
    reactants (2700 / 1396); 0, 1 or 2 products (1343 / 1348 / 1405);
    stoichiometry 1 or 2; K in [1e-8, 1]."""
    m = copasi
    assert m.S == 4096 and m.R == 4096 and len(m.k) == 4096 and not m.warnings
    L, Rm = m.dense()
    nreact = (L > 0).sum(axis=1)
    nprod = (Rm > 0).sum(axis=1)
    assert np.bincount(nreact).tolist()[1:] == [2700, 1396]
    assert np.bincount(nprod).tolist() == [1343, 1348, 1405]
    assert set(np.unique(L[L > 0])) <= {1, 2} and set(np.unique(Rm[Rm > 0])) <= {1, 2}
    k = np.array(m.k)
    assert k.min() >= 1e-8 and k.max() <= 1.0
    assert m.names[:3] == ["X0_in_compartment01", "X1_in_compartment01", "X2_in_compartment01"]


def test_reference_4096_rates_equal_the_kinetic_laws(copasi):
    """Each product-form formula ("K0* X1945* X3793") evaluated at x0 equals
    the converted mass-action rate k * prod x^L.  The 1365 homodimer laws
    are written as stochastic propensities ("K8* (1/2) * X1915* (X1915-1)");
    the converter, like the reference's, takes only their parameter and the
    stoichiometry (2 X -> ...), not the formula."""
    with gzip.open(FIXTURE, "rt") as f:
        text = f.read()
    formulas = re.findall(r'<kineticLaw formula="([^"]*)"', text)
    assert len(formulas) == 4096
    m = copasi
    short = [n.replace("_in_compartment01", "") for n in m.names]
    x = dict(zip(short, m.x0))
    col = {n: i for i, n in enumerate(short)}
    L, _ = m.dense()
    rates = np.array(m.k) * np.prod(np.array(m.x0)[None, :] ** L.astype(float), axis=1)
    homo = 0
    for r, fml in enumerate(formulas):
        dimer = re.fullmatch(r"K\d+\* \(1/2\) \* (X\d+)\* \(\1-1\)", fml)
        if dimer:
            homo += 1
            assert L[r, col[dimer.group(1)]] == 2 and L[r].sum() == 2
            continue
        toks = [t.strip() for t in fml.split("*")]
        v = m.k[r]
        for tok in toks[1:]:
            v *= x[tok]
        assert v == pytest.approx(rates[r], rel=1e-12)
    assert homo == 1365


def test_converted_folder_reads_back_and_evaluates(copasi, tmp_path):
    """The folder parses back to the same network (the LASSIE reader's
    formats) and the oracle's RHS of the converted model equals H (k prod x^L)."""
    sbml.write_biosimware(copasi, str(tmp_path), t=np.linspace(0, 1, 3))
    L = np.loadtxt(tmp_path / "left_side", dtype=int, delimiter="\t")
    Rm = np.loadtxt(tmp_path / "right_side", dtype=int, delimiter="\t")
    k = np.loadtxt(tmp_path / "c_vector")
    x0 = np.loadtxt(tmp_path / "M_0", delimiter="\t")
    L0, R0 = copasi.dense()
    assert np.array_equal(L, L0) and np.array_equal(Rm, R0)
    np.testing.assert_allclose(k, copasi.k, rtol=1e-6)   # "%e": 7 significant digits
    model = sbml.to_model(copasi)
    f = O.rhs(model, model.x0)
    rates = model.k * np.prod(model.x0[None, :] ** L0.astype(float), axis=1)
    np.testing.assert_allclose(f, (R0.astype(float) - L0).T @ rates, rtol=1e-12, atol=1e-300)
    assert x0.shape == (4096,)


def test_command_line(tmp_path):
    doc = tmp_path / "toy.sbml"
    doc.write_text(L2_DOC)
    assert sbml.main([str(doc), str(tmp_path / "out"), "--t-end", "10", "--samples", "5"]) == 0
    assert (tmp_path / "out" / "t_vector").read_text().split() == ["0", "2.5", "5", "7.5", "10"]
    assert sbml.main([str(tmp_path / "missing.sbml"), str(tmp_path / "o2")]) == 2

