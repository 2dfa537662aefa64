"""Small hand-built models shared by the tests (reaction lists -> dense
stoichiometry -> the same CSR the LASSIE parser/sparse build produce)."""
import numpy as np

from lassie_amd import synth


def from_reactions(S, reactions, k, x0, feed=None, atol=None):
    """reactions: list of (reactants, products) as {species: stoich} dicts."""
    R = len(reactions)
    L = np.zeros((R, S), np.int16)
    Rm = np.zeros((R, S), np.int16)
    for r, (lhs, rhs) in enumerate(reactions):
        for s, v in lhs.items():
            L[r, s] = v
        for s, v in rhs.items():
            Rm[r, s] = v
    return synth.model_from_dense(L, Rm, np.asarray(k, float), np.asarray(x0, float), feed, atol)


def decay(k=0.7, x0=1.0):
    """A -> 0."""
    return from_reactions(1, [({0: 1}, {})], [k], [x0])


def reversible(kf=2.0, kb=0.5, a0=1.0, b0=0.0):
    """A <-> B."""
    return from_reactions(2, [({0: 1}, {1: 1}), ({1: 1}, {0: 1})], [kf, kb], [a0, b0])


def chain(k1=1.0, k2=0.3):
    """A -> B -> C."""
    return from_reactions(3, [({0: 1}, {1: 1}), ({1: 1}, {2: 1})], [k1, k2], [1.0, 0.0, 0.0])


def dimer(k=3.0, a0=1.0):
    """2A -> B."""
    return from_reactions(2, [({0: 2}, {1: 1})], [k], [a0, 0.0])


def mixed():
    """zero-order source, bimolecular, homodimer, stoichiometry 2 product,
    and a spectator species: exercises every RHS / Jacobian branch."""
    reactions = [
        ({}, {0: 1}),               # 0 -> A
        ({0: 1, 1: 1}, {2: 1}),     # A + B -> C
        ({2: 2}, {3: 1}),           # 2C -> D
        ({3: 1}, {0: 1, 1: 2}),     # D -> A + 2B
        ({1: 1}, {}),               # B -> 0
        ({4: 1, 0: 1}, {4: 1, 2: 1}),  # E + A -> E + C (catalysed)
    ]
    return from_reactions(5, reactions, [0.5, 2.0, 1.5, 0.8, 0.1, 0.7],
                          [0.3, 0.9, 0.1, 0.05, 0.4])


def autocatalytic_singular(h=0.01):
    """A -> 2A with k = 1/h: I - hJ = 1 - hk = 0 at the first BDF step."""
    return from_reactions(1, [({0: 1}, {0: 2})], [1.0 / h], [1.0])


def with_feed():
    """A + B -> C with B held at a feed value."""
    return from_reactions(3, [({0: 1, 1: 1}, {2: 1}), ({2: 1}, {0: 1})], [1.2, 0.4],
                          [1.0, 0.5, 0.0], feed=[0.0, 0.5, 0.0])


def hub(partners=20):
    """A + B_k <-> C_k for k = 1..partners: the Jacobian row of A has
    2 partners + 1 distinct columns (more than the register-accumulated
    row form holds), the other rows three."""
    S = 1 + 2 * partners
    reactions, k = [], []
    for i in range(partners):
        b, c = 1 + i, 1 + partners + i
        reactions.append(({0: 1, b: 1}, {c: 1}))
        k.append(200.0 * (1 + i))
        reactions.append(({c: 1}, {0: 1, b: 1}))
        k.append(5.0 + i)
    x0 = [1.0] + [0.5] * partners + [0.0] * partners
    return from_reactions(S, reactions, k, x0)


def overflowing_rate(S=300, seed=11):
    """The stiff generator's S x S model with one bimolecular reaction of a
    species in the first 16 given k = 1e308 and both reactants 4: its
    Jacobian entries overflow to inf, so the Newton matrix has infinite
    entries in the first panel (pivot rows carrying inf, multipliers inf/inf)
    and every BDF step ends in a NaN repair."""
    m = synth.stiff(S, S, seed)
    L, Rm, k, x0 = m.L.copy(), m.Rm.copy(), m.k.copy(), m.x0.copy()
    r = next(i for i in range(m.R) if L[i].sum() == 2 and np.nonzero(L[i])[0][0] < 16)
    a, b = np.nonzero(L[r])[0]
    k[r] = 1e308
    x0[a] = x0[b] = 4.0
    return synth.model_from_dense(L, Rm, k, x0)
