"""LSODAsim: the reference's CPU LSODA simulator for BioSimWare folders.

Command line, files and numbers follow ``CPU-based LSODA/LSODAsim.py``
(main, 22-131) with ``src/GenerateODEs.py`` (open_files, 63-331) and
``src/LSODA.py`` (LSODER.run, 41-48):

    python -m lassie_amd.lsodasim input_folder output_folder [-v] [-f]

* **Inputs** (GenerateODEs.py:75-183): ``left_side``, ``right_side``,
  ``c_vector``, ``M_0``, ``t_vector``; optional ``M_feed``, ``alphabet``,
  ``cs_vector``, ``modelkind`` / ``volume`` (stochastic models: amounts and
  feeds divided by V * N_A, rate constants scaled by ``convertion_constant``,
  45-62), ``max_steps`` (default 10000), ``atol_vector`` (default 1e-12 per
  species), ``rtol`` (default 1e-6).
* **Equations** (GenerateODEs.py:203-331): mass action
  dy_s/dt = sum_r (R - L)[r, s] k_r prod y^L with the generated code's
  rules: a reactant stoichiometry other than 1 or 2 contributes no factor
  (the generator compares the file's text with '1' and '2'); a species with a
  nonzero feed has dy/dt = 0 and starts at its feed value when the feed is
  positive.  The analytic Jacobian of the same terms.
* **Solver** (LSODA.py:44): ``scipy.integrate.odeint`` with ``Dfun``,
  ``mxhnil=0``, ``mxstep=max_steps``, per-species ``atol``, ``rtol``.
* **Sampling** (LSODAsim.py:65-97): t = 0 is prepended when the first
  sampling time is not 0 and dropped from the output again; columns are the
  species of ``cs_vector``.
* **Outputs** (LSODAsim.py:99-124): ``output_folder/ODESol``, one row per
  sampling time (time, then the selected species) in "%.8e", tab-separated;
  with ``-f``, no file: the relative distance to ``input_folder/ts_matrix``
  (sum over nonzero target entries of |target - sim| / target, both floored
  at 1e-16) is printed instead.

This is the CPU tool the reference ships beside LASSIE; it does not use the
GPU.  The generated-Python-source approach of the reference is replaced by
vectorised numpy over the same stoichiometry.
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np
from scipy.integrate import odeint

N_A = 6.022e23


def _read_matrix(path):
    rows = []
    with open(path) as f:
        for line in f:
            line = line.strip("\n")
            if line.strip() == "":
                continue
            rows.append([int(float(tok)) if tok.strip() else 0 for tok in line.split("\t")])
    return np.array(rows, dtype=np.int64)


def _first_line_floats(path):
    with open(path) as f:
        return [float(tok) for tok in f.readline().strip("\n").split()]


def load_folder(folder: str) -> dict:
    """The model as GenerateODEs.open_files reads it (63-183)."""
    p = lambda name: os.path.join(folder, name)  # noqa: E731
    conv = 0
    try:
        with open(p("modelkind")) as f:
            kind = f.readline().split("\n")[0]
        conv = 1 if kind == "stochastic" else 0
    except OSError:
        conv = 0
    if conv:
        with open(p("volume")) as f:
            c_mm = float(f.readline()) * N_A
    else:
        c_mm = 1.0
    L = _read_matrix(p("left_side"))
    Rm = _read_matrix(p("right_side"))
    with open(p("c_vector")) as f:
        k = np.array([float(line) for line in f if line.strip()], float)
    x0 = np.array(_first_line_floats(p("M_0")), float)
    S = L.shape[1]
    try:
        feed = np.array(_first_line_floats(p("M_feed")), float)
    except OSError:
        feed = np.zeros(S)
    t = np.atleast_1d(np.loadtxt(p("t_vector")))
    try:
        with open(p("cs_vector")) as f:
            cs = [int(line) for line in f if line.strip()]
    except OSError:
        cs = list(range(S))
    x0 = x0 / c_mm
    feed = feed / c_mm
    x0 = np.where(feed > 0, feed, x0)
    if conv:
        k = convert_constants(k, L, c_mm)
    try:
        with open(p("max_steps")) as f:
            max_steps = int(f.readline())
    except OSError:
        max_steps = 10000
    try:
        atol = np.atleast_1d(np.loadtxt(p("atol_vector")))
    except OSError:
        atol = np.full(S, 1e-12)
    try:
        rtol = float(np.loadtxt(p("rtol")))
    except OSError:
        rtol = 1e-6
    return dict(L=L, Rm=Rm, k=k, x0=x0, feed=feed, t=t, cs=cs, atol=atol, rtol=rtol,
                max_steps=max_steps, S=S, R=L.shape[0])


def convert_constants(k, L, c_mm):
    """GenerateODEs.convertion_constant (45-62): homodimer k * c / 2,
    other bimolecular k * c, first order unchanged; order >= 3 is an error."""
    out = np.array(k, float)
    for r in range(L.shape[0]):
        if L[r].sum() >= 3:
            raise ValueError("There are one or more reactions with higher order to the second")
        if np.any(L[r] >= 2):
            out[r] = (k[r] * c_mm) / 2.0
        elif L[r].sum() > 1:
            out[r] = k[r] * c_mm
    return out


class GeneratedSystem:
    """dy/dt and its Jacobian as the generated ODE/JAC functions compute them."""

    def __init__(self, L, Rm, k, feed):
        self.D = (Rm - L).astype(float)            # R x S net change
        E = np.where((L == 1) | (L == 2), L, 0)    # only '1' and '2' make factors
        self.E = E.astype(float)
        self.k = np.asarray(k, float)
        self.frozen = np.asarray(feed) != 0
        self.rr, self.cc = np.nonzero(E)           # reactant factors

    def rates(self, y):
        v = self.k.copy()
        np.multiply.at(v, self.rr, y[self.cc] ** self.E[self.rr, self.cc])
        return v

    def f(self, y, t=0.0):
        dy = self.D.T @ self.rates(y)
        dy[self.frozen] = 0.0
        return dy

    def jac(self, y, t=0.0):
        R, S = self.D.shape
        dv = np.zeros((R, S))                      # d rate_r / d y_m
        for r, m in zip(self.rr, self.cc):
            e = self.E[r, m]
            d = self.k[r] * e * y[m] ** (e - 1)
            for r2, m2 in zip(self.rr, self.cc):
                if r2 == r and m2 != m:
                    d *= y[m2] ** self.E[r2, m2]
            dv[r, m] = d
        J = self.D.T @ dv
        J[self.frozen, :] = 0.0
        return J


class FastGeneratedSystem(GeneratedSystem):
    """The same Jacobian with the per-reaction products vectorised."""

    def __init__(self, L, Rm, k, feed):
        super().__init__(L, Rm, k, feed)
        R = self.D.shape[0]
        self.first = np.full(R, -1)
        self.second = np.full(R, -1)
        for r, m in zip(self.rr, self.cc):
            if self.first[r] < 0:
                self.first[r] = m
            elif self.second[r] < 0:
                self.second[r] = m
            else:
                self.first[r] = -2                 # three or more reactant species
        self.simple = bool(np.all(self.first >= -1))

    def jac(self, y, t=0.0):
        if not self.simple:
            return super().jac(y, t)
        R, S = self.D.shape
        dv = np.zeros((R, S))
        for a, b in ((self.first, self.second), (self.second, self.first)):
            rows = np.nonzero(a >= 0)[0]
            m = a[rows]
            e = self.E[rows, m]
            d = self.k[rows] * e * y[m] ** (e - 1)
            other = b[rows]
            has = other >= 0
            d[has] *= y[other[has]] ** self.E[rows[has], other[has]]
            dv[rows, m] = d
        J = self.D.T @ dv
        J[self.frozen, :] = 0.0
        return J


def simulate(model: dict):
    """odeint at the sampling times (LSODA.py:44, LSODAsim.py:65-97); returns
    the ODESol matrix: time, then the cs_vector species."""
    sysm = FastGeneratedSystem(model["L"], model["Rm"], model["k"], model["feed"])
    t = model["t"]
    prepend = t[0] != 0
    tt = np.concatenate([[0.0], t]) if prepend else t
    y = odeint(sysm.f, model["x0"], tt, Dfun=sysm.jac, mxhnil=0, printmessg=False,
               mxstep=model["max_steps"], atol=model["atol"], rtol=model["rtol"])
    if prepend:
        y = y[1:]
    out = np.zeros((t.size, len(model["cs"]) + 1))
    out[:, 0] = t
    out[:, 1:] = y[:, model["cs"]]
    return out


def fitness(sim, target):
    """LSODAsim.py:101-118: sum over nonzero target entries (columns from 1)
    of |target - sim| / target, both floored at 1e-16."""
    try:
        acc = 0.0
        for i in range(sim.shape[0]):
            for j in range(1, sim.shape[1]):
                vt = target[i, j]
                vs = sim[i, j]
                if vt != 0:
                    if target[i, j] < 1e-16:
                        vt = 1e-16
                    if sim[i, j] < 1e-16:
                        vs = 1e-16
                    acc += abs(vt - vs) / vt
        return acc
    except (IndexError, TypeError, ValueError):
        return sys.float_info.max


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    t0 = time.time()
    src = argv[0] if len(argv) > 0 else ""
    dst = argv[1] if len(argv) > 1 else ""
    dump = "-v" in argv[2:]
    fit = "-f" in argv[2:]
    target = None
    if fit:
        dump = False
        try:
            target = np.loadtxt(os.path.join(src, "ts_matrix"))
        except OSError:
            print(" * There is not ts_matrix file containing DTTS in", src)
    if dump:
        print(" * Opening project in folder", src)
        print(" * Saving output file in folder", dst)
    try:
        os.makedirs(dst)
    except OSError:
        if dump:
            print("WARNING:", dst, "already exists")
    model = load_folder(src)
    if dump:
        print("Stoichiometry size:", model["R"], "X", model["S"])
        print("Relative tolerance:", model["rtol"])
        print("Maximum number of steps:", model["max_steps"])
    sim = simulate(model)
    if fit:
        print(fitness(sim, target) if target is not None else sys.float_info.max)
    else:
        np.savetxt(os.path.join(dst, "ODESol"), sim, fmt="%.8e", delimiter="\t")
    if dump:
        print("* Dynamics saved in folder", os.path.join(dst, "ODESol"))
        print("Simulation time:", time.time() - t0)
    return 0


if __name__ == "__main__":
    sys.exit(main())
