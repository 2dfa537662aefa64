"""SBML (mass-action) -> BioSimWare input folder, without libsbml.

Restates the reference's converter ``gui/SBML2BSW.py`` (SBMLloader,
lines 30-350) with the standard library's XML parser, for SBML levels 1-3:

* species, in document order (``get_initial_amounts``, 58-105): the name is
  the species name, else its id, with ``_in_<compartment>`` appended when a
  compartment is set; a repeated name keeps its first index; the initial
  value is ``initialAmount``, else ``initialConcentration``;
* feeds (``convert_to_biosimware``, 303-305): species with
  ``constant="true"`` get feed value 1, all others 0;
* reactions (``get_reactions``, 110-294): reactant and product
  stoichiometries fill one row of ``left_side`` / ``right_side``; the rate
  constants are the kinetic law's local parameters, in order, or else the
  global parameters named in the rate law (in the order they appear).  A
  reaction marked reversible, or whose law names two global parameters, also
  gets its reverse (products and reactants swapped) right after it, with the
  second parameter;
* output files and formats (287-321): ``M_0`` ("%e", tab-separated),
  ``M_feed`` ("%d"), ``alphabet`` (names, each followed by a tab),
  ``left_side`` / ``right_side`` ("%d", tab-separated rows), ``c_vector``
  ("%e", one per line) and an empty ``boundaries`` (flux bounds of FBA
  models, which this converter does not read).

Not restated: FBA objectives and flux bounds (the ``fbc`` package), assignment
rules for parameters whose value is 0, and non-mass-action laws (the
reference's TODO list, 24-29).  ``t_vector`` is not part of the SBML; pass
``t`` to :func:`write_biosimware` (or ``--t-end/--samples`` on the command
line) to write one.

Usage::

    python -m lassie_amd.sbml model.sbml out_folder [--t-end 100 --samples 200]
"""
from __future__ import annotations

import argparse
import gzip
import math
import os
import re
import sys
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field

import numpy as np


class SBMLError(ValueError):
    pass


@dataclass
class SBMLModel:
    names: list = field(default_factory=list)       # species names (alphabet order = index)
    x0: list = field(default_factory=list)
    feed: list = field(default_factory=list)        # 1 for constant species
    left: list = field(default_factory=list)        # reactant stoichiometry rows
    right: list = field(default_factory=list)       # product stoichiometry rows
    k: list = field(default_factory=list)           # one rate constant per row
    reactions: list = field(default_factory=list)   # row names
    warnings: list = field(default_factory=list)

    @property
    def S(self):
        return len(self.names)

    @property
    def R(self):
        return len(self.left)

    def dense(self):
        """(L, Rm) as int16 R x S arrays."""
        L = np.array(self.left, dtype=np.int16).reshape(self.R, self.S)
        Rm = np.array(self.right, dtype=np.int16).reshape(self.R, self.S)
        return L, Rm


def _local(tag):
    return tag.rsplit("}", 1)[-1]


def _children(el, name):
    return [c for c in el if _local(c.tag) == name]


def _child(el, name):
    c = _children(el, name)
    return c[0] if c else None


def _fnum(s, default=math.nan):
    if s is None or s == "":
        return default
    return float(s)


def _law_tokens(law):
    """Identifiers used by a kinetic law, in order: the L1 ``formula``
    attribute tokenised as the reference does (re.findall(r"[\\w']+")), or
    the ``<ci>`` names of its MathML."""
    formula = law.get("formula")
    if formula is not None:
        return [t.strip() for t in re.findall(r"[\w']+", formula)]
    return [(ci.text or "").strip() for ci in law.iter() if _local(ci.tag) == "ci"]


def parse(source) -> SBMLModel:
    """Parse an SBML document (a path, optionally .gz, or the XML text)."""
    if isinstance(source, str) and source.lstrip().startswith("<"):
        root = ET.fromstring(source)
    else:
        opener = gzip.open if str(source).endswith(".gz") else open
        with opener(source, "rb") as f:
            root = ET.parse(f).getroot()
    if _local(root.tag) != "sbml":
        raise SBMLError("not an SBML document")
    model = _child(root, "model")
    if model is None:
        raise SBMLError("empty model or file not valid")
    out = SBMLModel()

    # species (SBML2BSW.py:58-105, id2name 325-335)
    index, id2idx = {}, {}
    los = _child(model, "listOfSpecies")
    species = _children(los, "species") if los is not None else []
    for sp in species:
        sid = sp.get("id") or sp.get("name") or ""
        name = sp.get("name") or sp.get("id") or ""
        comp = sp.get("compartment") or ""
        if comp:
            name = name + "_in_" + comp
        amount = _fnum(sp.get("initialAmount"))
        if math.isnan(amount):
            amount = _fnum(sp.get("initialConcentration"), 0.0)
        if name not in index:
            index[name] = len(out.names)
            out.names.append(name)
            out.x0.append(amount)
            out.feed.append(1 if (sp.get("constant") or "").lower() == "true" else 0)
        id2idx[sid] = index[name]
    S = len(out.names)

    # global parameters (SBML2BSW.py:140-154)
    glob = {}
    lop = _child(model, "listOfParameters")
    for p in (_children(lop, "parameter") if lop is not None else []):
        pid = p.get("id") or p.get("name")
        glob[pid] = _fnum(p.get("value"), 0.0)

    # reactions (SBML2BSW.py:166-294)
    lor = _child(model, "listOfReactions")
    for re_ in (_children(lor, "reaction") if lor is not None else []):
        rname = re_.get("id") or re_.get("name") or ""
        reversible = (re_.get("reversible") or "").lower() == "true"
        create_reverse = False
        params = []
        law = _child(re_, "kineticLaw")
        if law is not None:
            lp = _child(law, "listOfParameters") or _child(law, "listOfLocalParameters")
            local = []
            if lp is not None:
                local = [c for c in lp if _local(c.tag) in ("parameter", "localParameter")]
            if local:
                params = [_fnum(p.get("value"), 0.0) for p in local]
            else:
                used = [t for t in _law_tokens(law) if t in glob]
                if not used:
                    raise SBMLError(f"can't find any kinetic parameters for reaction {rname}")
                if reversible:
                    create_reverse = True
                elif len(used) == 2:
                    out.warnings.append(f"detected two parameters in kinetic law of reaction {rname}, "
                                        "assuming reversible reaction")
                    create_reverse = True
                elif len(used) > 2:
                    raise SBMLError("too many parameters in kinetic law")
                params = [glob[t] for t in used]
        left = [0] * S
        right = [0] * S
        for ref_list, vec in (("listOfReactants", left), ("listOfProducts", right)):
            lst = _child(re_, ref_list)
            for sr in (_children(lst, "speciesReference") if lst is not None else []):
                sid = sr.get("species")
                if sid not in id2idx:
                    raise SBMLError(f"reaction {rname} names unknown species {sid}")
                st = _fnum(sr.get("stoichiometry"), 1.0)
                vec[id2idx[sid]] = st
        out.k.extend(params)
        out.left.append([int(v) for v in left])
        out.right.append([int(v) for v in right])
        out.reactions.append(rname)
        if create_reverse:
            out.left.append([int(v) for v in right])
            out.right.append([int(v) for v in left])
            out.reactions.append(rname + " (reverse)")
    if len(out.k) != out.R:
        out.warnings.append(f"{len(out.k)} rate constants for {out.R} reactions")
    return out


def write_biosimware(m: SBMLModel, folder: str, t=None):
    """Write the BioSimWare files of SBML2BSW.convert_to_biosimware (287-321);
    with ``t``, also a ``t_vector``."""
    os.makedirs(folder, exist_ok=True)

    def path(name):
        return os.path.join(folder, name)

    with open(path("M_0"), "w") as f:
        f.write("\t".join("%e" % v for v in m.x0) + "\n")
    with open(path("M_feed"), "w") as f:
        f.write("\t".join("%d" % v for v in m.feed) + "\n")
    with open(path("alphabet"), "w") as f:
        f.write("".join(n + "\t" for n in m.names))
    for name, rows in (("left_side", m.left), ("right_side", m.right)):
        with open(path(name), "w") as f:
            f.write("".join("\t".join("%d" % v for v in row) + "\n" for row in rows))
    with open(path("c_vector"), "w") as f:
        f.write("".join("%e\n" % v for v in m.k))
    open(path("boundaries"), "w").close()
    if t is not None:
        with open(path("t_vector"), "w") as f:
            f.write("".join("%.17g\n" % v for v in t))


def to_model(m: SBMLModel):
    """The parsed network as the package's model record (synth.Model)."""
    from lassie_amd import synth
    L, Rm = m.dense()
    feed = np.array(m.feed, float)
    return synth.model_from_dense(L, Rm, np.array(m.k, float), np.array(m.x0, float),
                                  feed if feed.any() else None)


def main(argv=None):
    ap = argparse.ArgumentParser(description="SBML (mass action) to BioSimWare input folder")
    ap.add_argument("sbml")
    ap.add_argument("out")
    ap.add_argument("--t-end", type=float, default=None, help="also write t_vector = linspace(0, t_end, samples)")
    ap.add_argument("--samples", type=int, default=100)
    a = ap.parse_args(argv)
    try:
        m = parse(a.sbml)
    except (SBMLError, ET.ParseError, OSError) as e:
        print(f"ERROR: {e}", file=sys.stderr)
        return 2
    t = np.linspace(0.0, a.t_end, a.samples) if a.t_end is not None else None
    write_biosimware(m, a.out, t)
    for w in m.warnings:
        print("WARNING:", w)
    print(f" * {m.S} species, {m.R} reactions written to {a.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
