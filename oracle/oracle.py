"""ctypes front end of the oracle (``lassie_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py``, as the checker.  The product never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liblassie_oracle.so")

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class OrcModel(C.Structure):
    _fields_ = [("S", C.c_int32), ("R", C.c_int32),
                ("lptr", _i32p), ("lsp", _i32p), ("lexp", _i32p),
                ("hptr", _i32p), ("hrx", _i32p), ("hcoef", _i32p),
                ("k", _f64p), ("x0", _f64p), ("feed", _f64p), ("atol", _f64p)]


class OrcOpts(C.Structure):
    _fields_ = [("rkf_step", C.c_double), ("be_step", C.c_double), ("newton_tol", C.c_double),
                ("stiffness_tol", C.c_double), ("rtol", C.c_double), ("newton_iter", C.c_int32),
                ("bdf_regrow", C.c_int32), ("max_rkf_retries", C.c_int32), ("pad_", C.c_int32),
                ("max_iters", C.c_int64), ("refine_tol", C.c_double), ("refine", C.c_int32),
                ("refine_max_iter", C.c_int32), ("max_newton", C.c_int64)]


class OrcStats(C.Structure):
    _fields_ = [("loop_iters", C.c_int64), ("rkf_attempts", C.c_int64),
                ("rkf_accepts", C.c_int64), ("bdf_steps", C.c_int64),
                ("newton_iters", C.c_int64), ("lu_failures", C.c_int64),
                ("nan_repairs", C.c_int64), ("samples", C.c_int64),
                ("t_final", C.c_double), ("dt_final", C.c_double), ("newdt_final", C.c_double),
                ("refine_iters", C.c_int64), ("refine_fallbacks", C.c_int64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class OrcTape(C.Structure):
    _fields_ = [("cap", C.c_int64), ("n", C.c_int64), ("dt", _f64p), ("kind", _i32p),
                ("sample_step", _i32p), ("overflow", C.c_int32), ("pad_", C.c_int32)]


STEP_RKF = -1
STEP_LUFAIL = -2


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def use_library(path):
    """Loads the oracle from another build of the same sources (the tests
    that plant a mistake in a copy of the oracle use it); None restores the
    default build.  Returns the previously loaded library handle."""
    global _lib
    prev = _lib
    _lib = None
    if path is not None:
        lib(path)
    return prev


def lib(path=None):
    global _lib
    if _lib is None:
        path = path or _LIB
        if path == _LIB and not os.path.exists(_LIB):
            build()
        L = C.CDLL(path)
        L.orc_rhs.argtypes = [C.POINTER(OrcModel), _f64p, _f64p]
        L.orc_jac.argtypes = [C.POINTER(OrcModel), _f64p, _f64p]
        L.orc_newton_matrix.argtypes = [C.c_int, C.c_double, _f64p]
        L.orc_lu.argtypes = [C.c_int, _f64p, _i32p]
        L.orc_lu.restype = C.c_int
        L.orc_lu_solve.argtypes = [C.c_int, _f64p, _i32p, _f64p]
        L.orc_normsq.argtypes = [C.c_int, _f64p]
        L.orc_normsq.restype = C.c_double
        L.orc_min_first.argtypes = [C.c_int, _f64p]
        L.orc_min_first.restype = C.c_double
        L.orc_rkf_attempt.argtypes = [C.POINTER(OrcModel), C.c_double, C.c_double, _f64p, _f64p,
                                      _f64p, _f64p]
        L.orc_rkf_attempt.restype = C.c_int
        L.orc_solve.argtypes = [C.POINTER(OrcModel), C.POINTER(OrcOpts), _f64p, C.c_int, _i32p,
                                C.c_int, _f64p, _f64p, _f64p, C.c_int64, C.POINTER(OrcStats)]
        L.orc_solve.restype = C.c_int
        L.orc_set_lu_variant.argtypes = [C.c_int]
        L.orc_set_lu_variant.restype = None
        L.orc_solve_rec.argtypes = L.orc_solve.argtypes + [C.POINTER(OrcTape)]
        L.orc_solve_rec.restype = C.c_int
        L.orc_replay.argtypes = [C.POINTER(OrcModel), C.POINTER(OrcOpts), C.POINTER(OrcTape), C.c_int, _f64p]
        L.orc_replay.restype = C.c_int
        L.orc_tangent.argtypes = [C.POINTER(OrcModel), C.POINTER(OrcOpts), C.POINTER(OrcTape), C.c_int, _f64p,
                                  _f64p, _f64p]
        L.orc_tangent.restype = C.c_int
        L.orc_adjoint.argtypes = [C.POINTER(OrcModel), C.POINTER(OrcOpts), C.POINTER(OrcTape), C.c_int, _f64p,
                                  _f64p, _f64p]
        L.orc_adjoint.restype = C.c_int
        _f32p = C.POINTER(C.c_float)
        L.orc32_rhs.argtypes = [C.POINTER(OrcModel), _f32p, _f32p]
        L.orc32_jac.argtypes = [C.POINTER(OrcModel), _f32p, _f32p]
        L.orc32_newton_matrix.argtypes = [C.c_int, C.c_float, _f32p]
        L.orc32_lu.argtypes = [C.c_int, _f32p, _i32p]
        L.orc32_lu.restype = C.c_int
        L.orc32_lu_solve.argtypes = [C.c_int, _f32p, _i32p, _f32p]
        L.orc32_normsq.argtypes = [C.c_int, _f32p]
        L.orc32_normsq.restype = C.c_float
        L.orc32_rkf_attempt.argtypes = [C.POINTER(OrcModel), C.c_double, C.c_float, _f32p, _f32p, _f32p, _f32p]
        L.orc32_rkf_attempt.restype = C.c_int
        L.orc32_solve.argtypes = L.orc_solve.argtypes
        L.orc_refine_solve.argtypes = [C.c_int, _f64p, _f32p, _i32p, _f64p, _f64p, _f64p, _f64p, _f32p,
                                       C.c_double, C.c_int, C.POINTER(C.c_int64)]
        L.orc_refine_solve.restype = C.c_int
        L.orc32_solve.restype = C.c_int
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


class _Bound:
    """Keeps the numpy arrays alive while the C struct points at them."""

    def __init__(self, m):
        self.arrs = dict(
            lptr=np.ascontiguousarray(m.lptr, np.int32), lsp=np.ascontiguousarray(m.lsp, np.int32),
            lexp=np.ascontiguousarray(m.lexp, np.int32), hptr=np.ascontiguousarray(m.hptr, np.int32),
            hrx=np.ascontiguousarray(m.hrx, np.int32), hcoef=np.ascontiguousarray(m.hcoef, np.int32),
            k=np.ascontiguousarray(m.k, np.float64), x0=np.ascontiguousarray(m.x0, np.float64),
            feed=np.ascontiguousarray(m.feed, np.float64),
            atol=np.ascontiguousarray(m.atol, np.float64))
        a = self.arrs
        # zero-length arrays still need a valid pointer
        for key in ("lsp", "lexp", "hrx", "hcoef"):
            if a[key].size == 0:
                a[key] = np.zeros(1, np.int32)
        self.s = OrcModel(m.S, m.R, _p(a["lptr"], _i32p), _p(a["lsp"], _i32p),
                          _p(a["lexp"], _i32p), _p(a["hptr"], _i32p), _p(a["hrx"], _i32p),
                          _p(a["hcoef"], _i32p), _p(a["k"], _f64p), _p(a["x0"], _f64p),
                          _p(a["feed"], _f64p), _p(a["atol"], _f64p))


def opts(rkf_step=1e-3, be_step=1e-2, newton_tol=1e-6, stiffness_tol=1e-6, rtol=0.0,
         newton_iter=10000, bdf_regrow=0, max_rkf_retries=0, max_iters=0, refine=0,
         refine_tol=1e-14, refine_max_iter=4, max_newton=0):
    """Controller options; refine=1 selects the precision form (FP32 LU of
    (float)M with FP64 refinement, orc_refine_solve)."""
    return OrcOpts(rkf_step, be_step, newton_tol, stiffness_tol, rtol, newton_iter, bdf_regrow,
                   max_rkf_retries, 0, max_iters, refine_tol, refine, refine_max_iter, max_newton)


def rhs(m, x):
    b = _Bound(m)
    x = np.ascontiguousarray(x, np.float64)
    f = np.empty(m.S)
    lib().orc_rhs(C.byref(b.s), _p(x, _f64p), _p(f, _f64p))
    return f


def jac(m, x):
    b = _Bound(m)
    x = np.ascontiguousarray(x, np.float64)
    J = np.empty((m.S, m.S))
    lib().orc_jac(C.byref(b.s), _p(x, _f64p), _p(J, _f64p))
    return J


def newton_matrix(J, h):
    M = np.array(J, dtype=np.float64, order="C", copy=True)
    lib().orc_newton_matrix(M.shape[0], h, _p(M, _f64p))
    return M


def lu(A):
    LU = np.array(A, dtype=np.float64, order="C", copy=True)
    n = LU.shape[0]
    piv = np.zeros(n, np.int32)
    info = lib().orc_lu(n, _p(LU, _f64p), _p(piv, _i32p))
    return LU, piv, info


def lu_solve(LU, piv, b):
    x = np.array(b, dtype=np.float64, copy=True)
    LU = np.ascontiguousarray(LU, np.float64)
    piv = np.ascontiguousarray(piv, np.int32)
    lib().orc_lu_solve(LU.shape[0], _p(LU, _f64p), _p(piv, _i32p), _p(x, _f64p))
    return x


def refine_solve(M, LU32, piv, rhs, tol=1e-14, max_it=4, floor=None):
    """orc_refine_solve: returns (x, r, converged, sweeps)."""
    fl = None if floor is None else np.ascontiguousarray(floor, np.float64)
    M = np.ascontiguousarray(M, np.float64)
    LU32 = np.ascontiguousarray(LU32, np.float32)
    piv = np.ascontiguousarray(piv, np.int32)
    rhs = np.ascontiguousarray(rhs, np.float64)
    n = M.shape[0]
    x, r, b32 = np.empty(n), np.empty(n), np.empty(n, np.float32)
    it = C.c_int64(0)
    ok = lib().orc_refine_solve(n, _p(M, _f64p), _p(LU32, _f32p), _p(piv, _i32p), _p(rhs, _f64p),
                                None if fl is None else _p(fl, _f64p), _p(x, _f64p), _p(r, _f64p),
                                _p(b32, _f32p), tol, max_it, C.byref(it))
    return x, r, bool(ok), it.value


def normsq(d):
    d = np.ascontiguousarray(d, np.float64)
    return lib().orc_normsq(d.size, _p(d, _f64p))


def min_first(d):
    d = np.ascontiguousarray(d, np.float64)
    return lib().orc_min_first(d.size, _p(d, _f64p))


def rkf_attempt(m, y1, dt, rtol=0.0):
    b = _Bound(m)
    y1 = np.ascontiguousarray(y1, np.float64)
    w1, w2, delta = np.empty(m.S), np.empty(m.S), np.empty(m.S)
    nok = lib().orc_rkf_attempt(C.byref(b.s), rtol, dt, _p(y1, _f64p), _p(w1, _f64p),
                                _p(w2, _f64p), _p(delta, _f64p))
    return nok, w1, w2, delta


def solve(m, t, cs=None, progress_cap=0, **kw):
    """Full trajectory; returns (out[nt, ncs], stats dict, rc, progress, y_final)."""
    b = _Bound(m)
    t = np.ascontiguousarray(t, np.float64)
    cs = np.arange(m.S, dtype=np.int32) if cs is None else np.ascontiguousarray(cs, np.int32)
    out = np.full((t.size, cs.size), np.nan)
    o = opts(**kw)
    st = OrcStats()
    prog = np.zeros(max(progress_cap, 1))
    yf = np.zeros(m.S)
    rc = lib().orc_solve(C.byref(b.s), C.byref(o), _p(t, _f64p), t.size, _p(cs, _i32p), cs.size,
                         _p(out, _f64p), _p(yf, _f64p),
                         _p(prog, _f64p) if progress_cap else None, progress_cap, C.byref(st))
    d = st.as_dict()
    n_prog = min(progress_cap, d["loop_iters"])
    return out, d, rc, prog[:n_prog], yf


_f32p = C.POINTER(C.c_float)


def rhs32(m, x):
    """generate_ode in single precision (Simulation<float>)."""
    b = _Bound(m)
    x = np.ascontiguousarray(x, np.float32)
    f = np.empty(m.S, np.float32)
    lib().orc32_rhs(C.byref(b.s), _p(x, _f32p), _p(f, _f32p))
    return f


def jac32(m, x):
    b = _Bound(m)
    x = np.ascontiguousarray(x, np.float32)
    J = np.empty((m.S, m.S), np.float32)
    lib().orc32_jac(C.byref(b.s), _p(x, _f32p), _p(J, _f32p))
    return J


def lu32(A):
    LU = np.array(A, dtype=np.float32, order="C", copy=True)
    n = LU.shape[0]
    piv = np.zeros(n, np.int32)
    info = lib().orc32_lu(n, _p(LU, _f32p), _p(piv, _i32p))
    return LU, piv, info


def lu_solve32(LU, piv, b):
    x = np.array(b, dtype=np.float32, copy=True)
    LU = np.ascontiguousarray(LU, np.float32)
    piv = np.ascontiguousarray(piv, np.int32)
    lib().orc32_lu_solve(LU.shape[0], _p(LU, _f32p), _p(piv, _i32p), _p(x, _f32p))
    return x


def normsq32(d):
    d = np.ascontiguousarray(d, np.float32)
    return lib().orc32_normsq(d.size, _p(d, _f32p))


def rkf_attempt32(m, y1, dt, rtol=0.0):
    b = _Bound(m)
    y1 = np.ascontiguousarray(y1, np.float32)
    w1, w2, delta = (np.empty(m.S, np.float32) for _ in range(3))
    nok = lib().orc32_rkf_attempt(C.byref(b.s), rtol, dt, _p(y1, _f32p), _p(w1, _f32p), _p(w2, _f32p),
                                  _p(delta, _f32p))
    return nok, w1, w2, delta


def solve32(m, t, cs=None, progress_cap=0, **kw):
    """Full trajectory of the single-precision loop (the reference's `-p
    float`); same returns as solve(), values are floats widened to double."""
    b = _Bound(m)
    t = np.ascontiguousarray(t, np.float64)
    cs = np.arange(m.S, dtype=np.int32) if cs is None else np.ascontiguousarray(cs, np.int32)
    out = np.full((t.size, cs.size), np.nan)
    o = opts(**kw)
    st = OrcStats()
    prog = np.zeros(max(progress_cap, 1))
    yf = np.zeros(m.S)
    rc = lib().orc32_solve(C.byref(b.s), C.byref(o), _p(t, _f64p), t.size, _p(cs, _i32p), cs.size,
                           _p(out, _f64p), _p(yf, _f64p),
                           _p(prog, _f64p) if progress_cap else None, progress_cap, C.byref(st))
    d = st.as_dict()
    n_prog = min(progress_cap, d["loop_iters"])
    return out, d, rc, prog[:n_prog], yf


def set_lu_variant(two_roundings: bool):
    """Switch the oracle's LU trailing update between fma (the specified
    arithmetic) and separately rounded multiply/subtract (another valid FP64
    LU, for measuring the FP64 noise floor of trajectories)."""
    lib().orc_set_lu_variant(1 if two_roundings else 0)


class Tape:
    """Steps of one recorded solve: kind[n] (STEP_RKF, STEP_LUFAIL or the
    number of Newton iterations of a BDF1 step), dt[n], and sample_step[i]
    (steps taken when sample i was stored)."""

    def __init__(self, cap, nt):
        self.dt = np.zeros(cap)
        self.kind = np.zeros(cap, np.int32)
        self.sample_step = np.zeros(nt, np.int32)
        self.s = OrcTape(cap, 0, _p(self.dt, _f64p), _p(self.kind, _i32p), _p(self.sample_step, _i32p), 0, 0)
        self.opts = None

    @property
    def n(self):
        return self.s.n

    def steps(self):
        return self.kind[:self.n].copy(), self.dt[:self.n].copy()


def solve_tape(m, t, cap=1 << 20, **kw):
    """Full trajectory of every species plus the recorded step sequence:
    returns (out[nt, S], stats, rc, tape)."""
    b = _Bound(m)
    t = np.ascontiguousarray(t, np.float64)
    cs = np.arange(m.S, dtype=np.int32)
    out = np.full((t.size, m.S), np.nan)
    o = opts(**kw)
    st = OrcStats()
    tape = Tape(cap, t.size)
    rc = lib().orc_solve_rec(C.byref(b.s), C.byref(o), _p(t, _f64p), t.size, _p(cs, _i32p), m.S,
                             _p(out, _f64p), None, None, 0, C.byref(st), C.byref(tape.s))
    if tape.s.overflow:
        raise RuntimeError("tape capacity exceeded")
    tape.opts = kw
    return out, st.as_dict(), rc, tape


def _with_params(m, k=None, x0=None):
    import dataclasses
    return dataclasses.replace(m, k=m.k if k is None else np.asarray(k, np.float64),
                               x0=m.x0 if x0 is None else np.asarray(x0, np.float64))


def replay(m, tape, nt, k=None, x0=None):
    """States at the samples, the recorded steps re-run with rate constants k
    and initial state x0 (defaults: the model's)."""
    mm = _with_params(m, k, x0)
    b = _Bound(mm)
    o = opts(**tape.opts)
    out = np.full((nt, m.S), np.nan)
    rc = lib().orc_replay(C.byref(b.s), C.byref(o), C.byref(tape.s), nt, _p(out, _f64p))
    if rc != 0:
        raise RuntimeError(f"replay failed ({rc})")
    return out


def tangent(m, tape, nt, dlogk, dx0=None):
    """Forward sensitivity of the sampled states along (dlogk, dx0)."""
    b = _Bound(m)
    o = opts(**tape.opts)
    dlogk = np.ascontiguousarray(dlogk, np.float64)
    dx0 = np.zeros(m.S) if dx0 is None else np.ascontiguousarray(dx0, np.float64)
    dout = np.zeros((nt, m.S))
    rc = lib().orc_tangent(C.byref(b.s), C.byref(o), C.byref(tape.s), nt, _p(dlogk, _f64p), _p(dx0, _f64p),
                           _p(dout, _f64p))
    if rc != 0:
        raise RuntimeError(f"tangent failed ({rc})")
    return dout


def adjoint(m, tape, gout):
    """Reverse pass: (dL/dlog k [R], dL/dx0 [S]) for gout[i, s] = dL/d(state s
    at sample i)."""
    b = _Bound(m)
    o = opts(**tape.opts)
    gout = np.ascontiguousarray(gout, np.float64)
    gk, gx = np.zeros(m.R), np.zeros(m.S)
    rc = lib().orc_adjoint(C.byref(b.s), C.byref(o), C.byref(tape.s), gout.shape[0], _p(gout, _f64p),
                           _p(gk, _f64p), _p(gx, _f64p))
    if rc != 0:
        raise RuntimeError(f"adjoint failed ({rc})")
    return gk, gx
