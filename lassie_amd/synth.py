"""Seeded synthetic mass-action models and BioSimWare folder I/O.

A model is the in-memory form of a BioSimWare input folder after LASSIE's
parser and sparse build (``Reader.h:38-124``, ``Simulation.cuh:1061-1130``):

* ``L`` / ``Rm``: reactant / product stoichiometry, R x S (``left_side`` /
  ``right_side``, one tab-separated line per reaction);
* reactant CSR, reaction-major, species ascending (``create_sparse2``,
  ``Simulation.cuh:1107-1130``): ``lptr[R+1]``, ``lsp``, ``lexp``;
* net-stoichiometry CSR H = (Rm - L)^T, species-major, reaction ascending
  (``createTransposte`` + ``create_sparse``, ``Simulation.cuh:1061-1105``):
  ``hptr[S+1]``, ``hrx``, ``hcoef``;
* ``k`` (``c_vector``), ``x0`` (``M_0``), ``feed`` (``M_feed``), ``atol``
  (``atol_vector``, default 1e-12, ``Simulation.cuh:317-325``).

Generator (one reaction at a time, all draws vectorised in this fixed order
from ``numpy.random.default_rng(seed)``): number of reactants U{1,2}, number
of products U{0,1,2}, two distinct reactant candidates, two distinct product
candidates, then log-uniform rate constants, then initial amounts.
Stoichiometry is 1, so every reaction has order <= 2.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Model:
    S: int
    R: int
    L: np.ndarray | None  # int16 R x S (None for CSR-only models)
    Rm: np.ndarray | None
    lptr: np.ndarray
    lsp: np.ndarray
    lexp: np.ndarray
    hptr: np.ndarray
    hrx: np.ndarray
    hcoef: np.ndarray
    k: np.ndarray
    x0: np.ndarray
    feed: np.ndarray
    atol: np.ndarray
    meta: dict = field(default_factory=dict)

    @property
    def nnz_l(self) -> int:
        return int(self.lptr[-1])

    @property
    def nnz_h(self) -> int:
        return int(self.hptr[-1])


def csr_from_dense(L: np.ndarray, Rm: np.ndarray):
    """Reactant CSR (reaction-major) and net CSR (species-major) from dense
    stoichiometry, in the entry order of ``create_sparse2`` / ``create_sparse``
    (``Simulation.cuh:1079-1130``): row-major scans, zero entries skipped."""
    L = np.asarray(L, dtype=np.int32)
    Rm = np.asarray(Rm, dtype=np.int32)
    R, S = L.shape
    rr, ss = np.nonzero(L)
    lptr = np.zeros(R + 1, dtype=np.int32)
    np.add.at(lptr, rr + 1, 1)
    lptr = np.cumsum(lptr).astype(np.int32)
    lsp = ss.astype(np.int32)
    lexp = L[rr, ss].astype(np.int32)
    H = (Rm - L).T  # S x R
    hs, hr = np.nonzero(H)
    hptr = np.zeros(S + 1, dtype=np.int32)
    np.add.at(hptr, hs + 1, 1)
    hptr = np.cumsum(hptr).astype(np.int32)
    hrx = hr.astype(np.int32)
    hcoef = H[hs, hr].astype(np.int32)
    return lptr, lsp, lexp, hptr, hrx, hcoef


def model_from_dense(L, Rm, k, x0, feed=None, atol=None, meta=None) -> Model:
    L = np.asarray(L, dtype=np.int16)
    Rm = np.asarray(Rm, dtype=np.int16)
    R, S = L.shape
    lptr, lsp, lexp, hptr, hrx, hcoef = csr_from_dense(L, Rm)
    k = np.ascontiguousarray(k, dtype=np.float64)
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    feed = np.zeros(S) if feed is None else np.ascontiguousarray(feed, dtype=np.float64)
    atol = np.full(S, 1e-12) if atol is None else np.ascontiguousarray(atol, dtype=np.float64)
    assert k.shape == (R,) and x0.shape == (S,) and feed.shape == (S,) and atol.shape == (S,)
    return Model(S, R, L, Rm, lptr, lsp, lexp, hptr, hrx, hcoef, k, x0, feed, atol, dict(meta or {}))


def random_model(S: int, R: int, seed: int, k_range=(1e-3, 1e1), x0_kind="uniform",
                 x0_range=(0.0, 1.0)) -> Model:
    """The seeded generator: 1-2 distinct reactants, 0-2 distinct products,
    stoichiometry 1, log-uniform k, x0 uniform or log-uniform."""
    if S < 2 or R < 1:
        raise ValueError("generator needs S >= 2 and R >= 1")
    rng = np.random.default_rng(seed)
    n_react = rng.integers(1, 3, size=R)
    n_prod = rng.integers(0, 3, size=R)
    a = rng.integers(0, S, size=R)
    b = rng.integers(0, S - 1, size=R)
    b = b + (b >= a)
    c = rng.integers(0, S, size=R)
    d = rng.integers(0, S - 1, size=R)
    d = d + (d >= c)
    lk0, lk1 = np.log10(k_range[0]), np.log10(k_range[1])
    k = 10.0 ** rng.uniform(lk0, lk1, size=R)
    if x0_kind == "uniform":
        x0 = rng.uniform(x0_range[0], x0_range[1], size=S)
    elif x0_kind == "loguniform":
        x0 = 10.0 ** rng.uniform(np.log10(x0_range[0]), np.log10(x0_range[1]), size=S)
    else:
        raise ValueError(x0_kind)
    L = np.zeros((R, S), dtype=np.int16)
    Rm = np.zeros((R, S), dtype=np.int16)
    r = np.arange(R)
    L[r, a] = 1
    two = n_react == 2
    L[r[two], b[two]] = 1
    p1 = n_prod >= 1
    Rm[r[p1], c[p1]] = 1
    p2 = n_prod == 2
    Rm[r[p2], d[p2]] = 1
    meta = dict(S=S, R=R, seed=seed, k_range=list(k_range), x0_kind=x0_kind,
                x0_range=list(x0_range))
    return model_from_dense(L, Rm, k, x0, meta=meta)


# --- the workloads, with their sizes -------------------------------------------------------

def small_nonstiff(seed: int = 0) -> Model:
    """32 species x 48 reactions, k logU[1e-3, 1e1], x0 U[0, 1]."""
    return random_model(32, 48, seed, (1e-3, 1e1), "uniform", (0.0, 1.0))


def stiff(S: int, R: int, seed: int, k_range=(1e-6, 1e6)) -> Model:
    """Stiff generator: k logU[k_range], x0 logU[1e-6, 1]."""
    return random_model(S, R, seed, k_range, "loguniform", (1e-6, 1.0))


def uneven_batch(n: int, seed_base: int = 1000, sizes=(16, 32, 64, 128, 256, 512),
                 k_range=(1e-4, 1e4)) -> list[Model]:
    """n independent instances, S drawn uniformly from ``sizes``, R = 2S,
    instance i generated from seed ``seed_base + i``."""
    rng = np.random.default_rng(seed_base)
    pick = rng.integers(0, len(sizes), size=n)
    return [stiff(int(sizes[p]), 2 * int(sizes[p]), seed_base + i, k_range)
            for i, p in enumerate(pick)]


def parameterisations(S: int, R: int, B: int, seed: int = 0, k_range=(1e-4, 1e4),
                      hidden_seed: int = 7) -> tuple[list[Model], np.ndarray]:
    """B parameterisations of one stiff network (structure and x0 from
    ``seed``): parameterisation b has its own log-uniform rates (seed
    ``seed + 1 + b``).  Also returns the hidden rates k* = k * U[0.5, 2]
    (seed ``hidden_seed``), [B, R], that generate a target trajectory."""
    base = stiff(S, R, seed, k_range)
    lk0, lk1 = np.log10(k_range[0]), np.log10(k_range[1])
    models, hidden = [], []
    hrng = np.random.default_rng(hidden_seed)
    for b in range(B):
        k = 10.0 ** np.random.default_rng(seed + 1 + b).uniform(lk0, lk1, size=R)
        models.append(model_from_dense(base.L, base.Rm, k, base.x0,
                                       meta=dict(base.meta, parameterisation=b)))
        hidden.append(k * hrng.uniform(0.5, 2.0, size=R))
    return models, np.array(hidden)


def t_vector(t0: float, t1: float, n: int) -> np.ndarray:
    """Equispaced sampling times, as the LASSIE GUI writes them
    (``gui/lassie-gui.py:341-344``)."""
    return np.linspace(t0, t1, n)


# --- BioSimWare folders ---------------------------------------------------------------------

def write_folder(m: Model, folder: str, t: np.ndarray, cs=None, extra: dict | None = None):
    """Write a BioSimWare input folder readable by ``LASSIE`` (file names from
    ``simulator.cu:81-152``, formats from ``Reader.h``)."""
    if m.L is None:
        raise ValueError("model has no dense stoichiometry")
    os.makedirs(folder, exist_ok=True)

    def mat(path, A):
        with open(path, "w") as f:
            for row in A:
                f.write("\t".join(str(int(v)) for v in row) + "\n")

    mat(os.path.join(folder, "left_side"), m.L)
    mat(os.path.join(folder, "right_side"), m.Rm)
    with open(os.path.join(folder, "c_vector"), "w") as f:
        f.write("".join(f"{v:.17g}\n" for v in m.k))
    with open(os.path.join(folder, "M_0"), "w") as f:
        f.write("\t".join(f"{v:.17g}" for v in m.x0) + "\n")
    with open(os.path.join(folder, "t_vector"), "w") as f:
        f.write("".join(f"{v:.17g}\n" for v in t))
    if np.any(m.feed != 0):
        with open(os.path.join(folder, "M_feed"), "w") as f:
            f.write("\t".join(f"{v:.17g}" for v in m.feed) + "\n")
    if np.any(m.atol != 1e-12):
        with open(os.path.join(folder, "atol_vector"), "w") as f:
            f.write("".join(f"{v:.17g}\n" for v in m.atol))
    if cs is not None:
        with open(os.path.join(folder, "cs_vector"), "w") as f:
            f.write("".join(f"{int(v)}\n" for v in cs))
    for name, val in (extra or {}).items():
        with open(os.path.join(folder, name), "w") as f:
            f.write(f"{val}\n")
