"""CPU LSODA path: the reference's CPU baseline restated with numpy/scipy.

TEST INFRASTRUCTURE ONLY: used by tests/ and by bench.py's cpu_baseline leg as
checker and CPU reference; the product never imports it.

Reference behaviour restated here (``CPU-based LSODA/``, Python 2, never
imported or run):
  * ``LSODER.run`` (``src/LSODA.py:41-48``): ``scipy.integrate.odeint`` with
    ``Dfun=JAC``, ``mxhnil=0``, ``mxstep=max_steps`` (default 10000,
    ``GenerateODEs.py:163-166``), per-species ``atol`` (default 1e-12,
    ``GenerateODEs.py:168-171``) and ``rtol`` (default 1e-6, ``:173-176``);
  * the ODE and Jacobian (``GenerateODEs.py:183-331``): mass action
    f = H (k * prod x^L); feed species have dy/dt = 0 and start at their feed
    value (``GenerateODEs.py:157-159, 220-223``);
  * ``LSODAsim.py:65-71``: t = 0 is prepended when the first sampling time
    is not 0, and dropped from the output again (``:85-97``).
The reference writes the ODE and Jacobian as generated Python source; here
they are vectorised numpy over the same CSR the GPU path uses (reactant
exponents of any order, where the generated code handles only 1 and 2,
``GenerateODEs.py:238-247``).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
from scipy.integrate import odeint


class MassAction:
    def __init__(self, m):
        self.S, self.R = m.S, m.R
        self.k = np.asarray(m.k, float)
        rr = np.repeat(np.arange(m.R), np.diff(m.lptr))
        self.rr = rr
        self.sp = np.asarray(m.lsp, int)
        self.ex = np.asarray(m.lexp, int)
        hs = np.repeat(np.arange(m.S), np.diff(m.hptr))
        self.H = sp.csr_matrix((np.asarray(m.hcoef, float), (hs, np.asarray(m.hrx, int))),
                               shape=(m.S, m.R))
        self.feed_mask = np.asarray(m.feed) != 0

    def rates(self, x):
        terms = x[self.sp] ** self.ex
        v = self.k.copy()
        np.multiply.at(v, self.rr, terms)
        return v

    def f(self, x, t=0.0):
        dx = self.H @ self.rates(x)
        dx[self.feed_mask] = 0.0
        return dx

    def jac(self, x, t=0.0):
        # d v_r / d x_g = k_r * e_g x_g^(e_g - 1) * prod_{other reactants} x^e
        n = self.rr.size
        vals = np.empty(n)
        terms = x[self.sp] ** self.ex
        for idx in range(n):
            r = self.rr[idx]
            g, e = self.sp[idx], self.ex[idx]
            d = self.k[r] * e * x[g] ** (e - 1)
            others = (self.rr == r)
            others[idx] = False
            vals[idx] = d * np.prod(terms[others])
        D = sp.csr_matrix((vals, (self.rr, self.sp)), shape=(self.R, self.S))
        J = (self.H @ D).toarray()
        J[self.feed_mask, :] = 0.0
        return J


class FastMassAction(MassAction):
    """Same functions with the Jacobian values vectorised for order <= 2."""

    def __init__(self, m):
        super().__init__(m)
        counts = np.diff(np.asarray(m.lptr))
        self.order_ok = counts.max(initial=0) <= 2
        first = np.asarray(m.lptr[:-1])
        self.partner = np.full(self.rr.size, -1)
        for r in np.nonzero(counts == 2)[0]:
            a = first[r]
            self.partner[a] = a + 1
            self.partner[a + 1] = a

    def jac(self, x, t=0.0):
        if not self.order_ok:
            return super().jac(x, t)
        terms = x[self.sp] ** self.ex
        d = self.k[self.rr] * self.ex * x[self.sp] ** (self.ex - 1)
        has = self.partner >= 0
        d[has] *= terms[self.partner[has]]
        D = sp.csr_matrix((d, (self.rr, self.sp)), shape=(self.R, self.S))
        J = (self.H @ D).toarray()
        J[self.feed_mask, :] = 0.0
        return J


def solve(m, t, cs=None, rtol=1e-6, atol=None, max_steps=10000):
    """Trajectory at the sampling times t: out[nt, ncs]."""
    sys_ = FastMassAction(m)
    y0 = np.array(m.x0, float)
    fm = np.asarray(m.feed) > 0
    y0[fm] = np.asarray(m.feed)[fm]
    atol = np.full(m.S, 1e-12) if atol is None else atol
    t = np.asarray(t, float)
    prepend = t[0] != 0
    tt = np.concatenate([[0.0], t]) if prepend else t
    y = odeint(sys_.f, y0, tt, Dfun=sys_.jac, mxhnil=0, mxstep=max_steps, atol=atol, rtol=rtol,
               printmessg=False)
    if prepend:
        y = y[1:]
    cs = np.arange(m.S) if cs is None else np.asarray(cs)
    return y[:, cs]
