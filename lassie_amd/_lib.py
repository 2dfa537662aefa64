"""ctypes binding of liblassie.so (the C ABI declared in include/lassie.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C lassie_amd/csrc``).  Nothing here falls back to another
implementation: if the library is missing, or no HIP device is visible when a
model is created, the call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblassie.so")

LASSIE_OK = 0
LASSIE_ERR_INVALID = 1
LASSIE_ERR_HIP = 2
LASSIE_ERR_NOMEM = 3
LASSIE_ERR_MAXITER = 4
LASSIE_ERR_UNSUPPORTED = 5
LASSIE_ERR_DEVICE = 6
LASSIE_ERR_IO = 7
LASSIE_ERR_NONFINITE = 8

PREC_FP64 = 0
PREC_FP32_REFINE = 1
PREC_BF16_REFINE = 2
PREC_FP32_LOOP = 3
BATCH_VARIANTS = 3

_i32p = C.POINTER(C.c_int32)
_i16p = C.POINTER(C.c_int16)
_f64p = C.POINTER(C.c_double)


class ModelDesc(C.Structure):
    _fields_ = [("n_species", C.c_int32), ("n_reactions", C.c_int32),
                ("react_ptr", _i32p), ("react_species", _i32p), ("react_exp", _i32p),
                ("net_ptr", _i32p), ("net_reaction", _i32p), ("net_coef", _i32p),
                ("k", _f64p), ("x0", _f64p), ("feed", _f64p), ("atol", _f64p)]


class Opts(C.Structure):
    _fields_ = [("rkf_step", C.c_double), ("be_step", C.c_double), ("newton_tol", C.c_double),
                ("stiffness_tol", C.c_double), ("rtol", C.c_double), ("newton_iter", C.c_int32),
                ("bdf_regrow", C.c_int32), ("max_rkf_retries", C.c_int32),
                ("precision", C.c_int32), ("max_iters", C.c_int64), ("refine_tol", C.c_double),
                ("refine_max_iter", C.c_int32), ("reserved", C.c_int32), ("max_newton", C.c_int64)]


class Stats(C.Structure):
    _fields_ = [("loop_iters", C.c_int64), ("rkf_attempts", C.c_int64),
                ("rkf_accepts", C.c_int64), ("bdf_steps", C.c_int64),
                ("newton_iters", C.c_int64), ("lu_failures", C.c_int64),
                ("nan_repairs", C.c_int64), ("samples", C.c_int64),
                ("t_final", C.c_double), ("dt_final", C.c_double), ("newdt_final", C.c_double),
                ("kernel_ms", C.c_double), ("lu_factorizations", C.c_int64),
                ("refine_iters", C.c_int64), ("refine_fallbacks", C.c_int64),
                ("phase_ms", C.c_double * 4), ("groups", C.c_int32), ("lds_model", C.c_int32)]

    def as_dict(self):
        d = {n: getattr(self, n) for n, _ in self._fields_}
        d["phase_ms"] = list(self.phase_ms)
        return d


# every symbol include/lassie.h declares
EXPORTS = ("lassie_opts_default", "lassie_model_create", "lassie_model_destroy", "lassie_solve",
           "lassie_solve_batch", "lassie_solve_batch_static", "lassie_model_set_rates", "lassie_forward_batch",
           "lassie_adjoint_batch", "lassie_adjoint_status", "lassie_tape_info", "lassie_tape_read", "lassie_tape_destroy",
           "lassie_batch_plan_create", "lassie_batch_plan_run", "lassie_batch_plan_info",
           "lassie_batch_plan_progress", "lassie_batch_plan_timing", "lassie_batch_plan_destroy",
           "lassie_csr_from_dense", "lassie_device_count", "lassie_last_error", "lassie_abi_version")


class LassieError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"liblassie error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                               "or make -C lassie_amd/csrc")
        L = C.CDLL(LIB_PATH)
        L.lassie_opts_default.argtypes = [C.POINTER(Opts)]
        L.lassie_opts_default.restype = None
        L.lassie_model_create.argtypes = [C.POINTER(ModelDesc), C.POINTER(C.c_void_p)]
        L.lassie_model_destroy.argtypes = [C.c_void_p]
        L.lassie_solve.argtypes = [C.c_void_p, C.POINTER(Opts), _f64p, C.c_int32, _i32p, C.c_int32,
                                   _f64p, _f64p, _f64p, C.c_int64, C.POINTER(Stats)]
        L.lassie_solve_batch.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(Opts), _f64p,
                                         C.c_int32, _f64p, _f64p, C.POINTER(Stats),
                                         C.POINTER(C.c_double)]
        L.lassie_solve_batch_static.argtypes = L.lassie_solve_batch.argtypes
        L.lassie_csr_from_dense.argtypes = [C.c_int32, C.c_int32, _i16p, _i16p, _i32p, _i32p,
                                            _i32p, _i32p, _i32p, _i32p, _i32p, _i32p]
        L.lassie_model_set_rates.argtypes = [C.c_void_p, _f64p]
        L.lassie_forward_batch.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(Opts), _f64p,
                                           C.c_int32, C.c_int64, _f64p, C.POINTER(Stats),
                                           C.POINTER(C.c_double), C.POINTER(C.c_void_p)]
        L.lassie_adjoint_batch.argtypes = [C.c_void_p, _f64p, _f64p, _f64p, C.POINTER(C.c_double)]
        L.lassie_tape_info.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int64), _i32p]
        L.lassie_adjoint_status.argtypes = [C.c_void_p, C.c_int32, _i32p, _i32p]
        L.lassie_tape_read.argtypes = [C.c_void_p, C.c_int32, _f64p, _i32p, _i32p]
        L.lassie_tape_destroy.argtypes = [C.c_void_p]
        L.lassie_batch_plan_create.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(Opts), _f64p,
                                               C.c_int32, C.POINTER(C.c_void_p)]
        L.lassie_batch_plan_run.argtypes = [C.c_void_p, _f64p, _f64p, C.POINTER(Stats), C.POINTER(C.c_double)]
        L.lassie_batch_plan_info.argtypes = [C.c_void_p, _i32p, _i32p, C.POINTER(C.c_float)]
        L.lassie_batch_plan_progress.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        L.lassie_batch_plan_timing.argtypes = [C.c_void_p, _i32p, _i32p, C.POINTER(C.c_float)]
        L.lassie_batch_plan_destroy.argtypes = [C.c_void_p]
        L.lassie_device_count.argtypes = [_i32p]
        L.lassie_last_error.restype = C.c_char_p
        L.lassie_abi_version.restype = C.c_int
        _lib = L
    return _lib


def _check(rc):
    if rc != LASSIE_OK:
        raise LassieError(rc, lib().lassie_last_error().decode())


def _p(a, t):
    return a.ctypes.data_as(t)


def default_opts(**kw) -> Opts:
    o = Opts()
    lib().lassie_opts_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def device_count() -> int:
    n = C.c_int32(0)
    rc = lib().lassie_device_count(C.byref(n))
    return n.value if rc == LASSIE_OK else 0


def csr_from_dense(L: np.ndarray, Rm: np.ndarray):
    """The library's own dense -> CSR build (create_sparse / create_sparse2)."""
    L = np.ascontiguousarray(L, np.int16)
    Rm = np.ascontiguousarray(Rm, np.int16)
    R, S = L.shape
    nl, nh = C.c_int32(0), C.c_int32(0)
    _check(lib().lassie_csr_from_dense(S, R, _p(L, _i16p), _p(Rm, _i16p), C.byref(nl), C.byref(nh),
                                       None, None, None, None, None, None))
    out = [np.zeros(R + 1, np.int32), np.zeros(max(nl.value, 1), np.int32),
           np.zeros(max(nl.value, 1), np.int32), np.zeros(S + 1, np.int32),
           np.zeros(max(nh.value, 1), np.int32), np.zeros(max(nh.value, 1), np.int32)]
    _check(lib().lassie_csr_from_dense(S, R, _p(L, _i16p), _p(Rm, _i16p), None, None,
                                       *[_p(a, _i32p) for a in out]))
    lptr, lsp, lexp, hptr, hrx, hcoef = out
    return lptr, lsp[:nl.value], lexp[:nl.value], hptr, hrx[:nh.value], hcoef[:nh.value]


class DeviceModel:
    """A model resident in device memory (lassie_model_create)."""

    def __init__(self, m):
        self.S, self.R = m.S, m.R
        keep = dict(
            lptr=np.ascontiguousarray(m.lptr, np.int32), lsp=np.ascontiguousarray(m.lsp, np.int32),
            lexp=np.ascontiguousarray(m.lexp, np.int32), hptr=np.ascontiguousarray(m.hptr, np.int32),
            hrx=np.ascontiguousarray(m.hrx, np.int32), hcoef=np.ascontiguousarray(m.hcoef, np.int32),
            k=np.ascontiguousarray(m.k, np.float64), x0=np.ascontiguousarray(m.x0, np.float64),
            feed=np.ascontiguousarray(m.feed, np.float64),
            atol=np.ascontiguousarray(m.atol, np.float64))
        for key in ("lsp", "lexp", "hrx", "hcoef"):
            if keep[key].size == 0:
                keep[key] = np.zeros(1, np.int32)
        d = ModelDesc(m.S, m.R, _p(keep["lptr"], _i32p), _p(keep["lsp"], _i32p),
                      _p(keep["lexp"], _i32p), _p(keep["hptr"], _i32p), _p(keep["hrx"], _i32p),
                      _p(keep["hcoef"], _i32p), _p(keep["k"], _f64p), _p(keep["x0"], _f64p),
                      _p(keep["feed"], _f64p), _p(keep["atol"], _f64p))
        h = C.c_void_p()
        _check(lib().lassie_model_create(C.byref(d), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().lassie_model_destroy(self.h)
            self.h = None

    def set_rates(self, k):
        """Replace the rate constants (lassie_model_set_rates)."""
        k = np.ascontiguousarray(k, np.float64)
        if k.shape != (self.R,):
            raise ValueError(f"k must have shape ({self.R},)")
        _check(lib().lassie_model_set_rates(self.h, _p(k, _f64p)))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, t, cs=None, progress_cap=0, raise_on_maxiter=True, **opts):
        """Returns (out[nt, ncs], y_final[S], stats dict, status, progress)."""
        t = np.ascontiguousarray(t, np.float64)
        cs = np.arange(self.S, dtype=np.int32) if cs is None else np.ascontiguousarray(cs, np.int32)
        out = np.full((t.size, cs.size), np.nan)
        yf = np.zeros(self.S)
        prog = np.zeros(max(progress_cap, 1))
        st = Stats()
        o = default_opts(**opts)
        rc = lib().lassie_solve(self.h, C.byref(o), _p(t, _f64p), t.size, _p(cs, _i32p), cs.size,
                                _p(out, _f64p), _p(yf, _f64p),
                                _p(prog, _f64p) if progress_cap else None, progress_cap,
                                C.byref(st))
        if rc != LASSIE_OK and not (rc == LASSIE_ERR_MAXITER and not raise_on_maxiter):
            _check(rc)
        d = st.as_dict()
        return out, yf, d, rc, prog[:min(progress_cap, d["loop_iters"])]


def solve_batch(models, t, raise_on_maxiter=True, static=False, **opts):
    """Batched form over uploaded DeviceModels.  Returns (list of out[nt, S_i],
    list of y_final[S_i], list of stats dicts, device ms, status).  static:
    the entry point that allocates no device memory
    (lassie_solve_batch_static)."""
    t = np.ascontiguousarray(t, np.float64)
    n = len(models)
    sizes = [dm.S for dm in models]
    out = np.full(t.size * sum(sizes), np.nan)
    yf = np.zeros(sum(sizes))
    st = (Stats * n)()
    hs = (C.c_void_p * n)(*[dm.h for dm in models])
    ms = C.c_double(0.0)
    o = default_opts(**opts)
    fn = lib().lassie_solve_batch_static if static else lib().lassie_solve_batch
    rc = fn(hs, n, C.byref(o), _p(t, _f64p), t.size, _p(out, _f64p), _p(yf, _f64p), st, C.byref(ms))
    if rc != LASSIE_OK and not (rc == LASSIE_ERR_MAXITER and not raise_on_maxiter):
        _check(rc)
    outs, ys = [], []
    o_off = y_off = 0
    for S in sizes:
        outs.append(out[o_off:o_off + t.size * S].reshape(t.size, S))
        ys.append(yf[y_off:y_off + S])
        o_off += t.size * S
        y_off += S
    return outs, ys, [s.as_dict() for s in st], ms.value, rc


class Tape:
    """Step record of one forward pass of the differentiation form
    (lassie_forward_batch); keeps its models alive until closed."""

    def __init__(self, h, models, nt):
        self.h = h
        self.models = list(models)
        self.nt = nt

    def close(self):
        if getattr(self, "h", None):
            lib().lassie_tape_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self, i):
        n = C.c_int64(0)
        mx = np.zeros(1, np.int32)
        _check(lib().lassie_tape_info(self.h, i, C.byref(n), _p(mx, _i32p)))
        return n.value, int(mx[0])

    def read(self, i):
        """(dt[n], kind[n], sample_step[nt]) of model i."""
        n, _ = self.info(i)
        dt = np.zeros(max(n, 1))
        kind = np.zeros(max(n, 1), np.int32)
        ss = np.zeros(self.nt, np.int32)
        _check(lib().lassie_tape_read(self.h, i, _p(dt, _f64p), _p(kind, _i32p), _p(ss, _i32p)))
        return dt[:n], kind[:n], ss

    def status(self, i):
        """(finite, capped_newton_steps) of model i: whether its last gradient
        was finite (1), not finite (0) or not computed yet (-1), and how many
        of its BDF1 steps ran the full newton_iter Newton iterations."""
        f = np.zeros(1, np.int32)
        c = np.zeros(1, np.int32)
        _check(lib().lassie_adjoint_status(self.h, i, _p(f, _i32p), _p(c, _i32p)))
        return int(f[0]), int(c[0])

    def adjoint(self, grad_outs, allow_nonfinite=False):
        """grad_outs: list of dL/d(out_i) [nt, S_i].  Returns (list of dL/dlog k
        [R_i], list of dL/dx0 [S_i], device ms).  A gradient that is not
        finite (a BDF1 step whose modified Newton iteration did not converge)
        raises LassieError unless allow_nonfinite; status(i) tells which."""
        g = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float64).ravel() for x in grad_outs]))
        gk = np.zeros(sum(m.R for m in self.models))
        gx = np.zeros(sum(m.S for m in self.models))
        ms = C.c_double(0.0)
        rc = lib().lassie_adjoint_batch(self.h, _p(g, _f64p), _p(gk, _f64p), _p(gx, _f64p), C.byref(ms))
        if rc != LASSIE_OK and not (rc == LASSIE_ERR_NONFINITE and allow_nonfinite):
            _check(rc)
        gks, gxs = [], []
        ko = xo = 0
        for m in self.models:
            gks.append(gk[ko:ko + m.R])
            gxs.append(gx[xo:xo + m.S])
            ko += m.R
            xo += m.S
        return gks, gxs, ms.value


def forward_batch(models, t, max_steps=1 << 16, **opts):
    """Forward pass of the differentiation form.  Returns (list of out[nt, S_i],
    list of stats dicts, device ms, Tape)."""
    t = np.ascontiguousarray(t, np.float64)
    n = len(models)
    sizes = [dm.S for dm in models]
    out = np.full(t.size * sum(sizes), np.nan)
    st = (Stats * n)()
    hs = (C.c_void_p * n)(*[dm.h for dm in models])
    ms = C.c_double(0.0)
    o = default_opts(**opts)
    th = C.c_void_p()
    _check(lib().lassie_forward_batch(hs, n, C.byref(o), _p(t, _f64p), t.size, max_steps, _p(out, _f64p), st,
                                      C.byref(ms), C.byref(th)))
    outs = []
    off = 0
    for S in sizes:
        outs.append(out[off:off + t.size * S].reshape(t.size, S))
        off += t.size * S
    return outs, [s.as_dict() for s in st], ms.value, Tape(th, models, t.size)


class BatchPlan:
    """The batched form prepared once and replayed as a graph
    (lassie_batch_plan_*); the kernel variant is chosen by the autotuner."""

    def __init__(self, models, t, **opts):
        self.models = list(models)
        self.t = np.ascontiguousarray(t, np.float64)
        n = len(self.models)
        hs = (C.c_void_p * n)(*[dm.h for dm in self.models])
        o = default_opts(**opts)
        h = C.c_void_p()
        _check(lib().lassie_batch_plan_create(hs, n, C.byref(o), _p(self.t, _f64p), self.t.size, C.byref(h)))
        self.h = h

    def run(self, raise_on_maxiter=True):
        """Returns (list of out[nt, S_i], list of y_final[S_i], list of stats, device ms, status)."""
        n = len(self.models)
        sizes = [dm.S for dm in self.models]
        out = np.full(self.t.size * sum(sizes), np.nan)
        yf = np.zeros(sum(sizes))
        st = (Stats * n)()
        ms = C.c_double(0.0)
        rc = lib().lassie_batch_plan_run(self.h, _p(out, _f64p), _p(yf, _f64p), st, C.byref(ms))
        if rc != LASSIE_OK and not (rc == LASSIE_ERR_MAXITER and not raise_on_maxiter):
            _check(rc)
        outs, ys = [], []
        o_off = y_off = 0
        for S in sizes:
            outs.append(out[o_off:o_off + self.t.size * S].reshape(self.t.size, S))
            ys.append(yf[y_off:y_off + S])
            o_off += self.t.size * S
            y_off += S
        return outs, ys, [x.as_dict() for x in st], ms.value, rc

    def info(self):
        """(variant, row tiles in flight, trial ms per variant)."""
        v = np.zeros(1, np.int32)
        rt = np.zeros(1, np.int32)
        ms = (C.c_float * BATCH_VARIANTS)()
        _check(lib().lassie_batch_plan_info(self.h, _p(v, _i32p), _p(rt, _i32p), ms))
        return int(v[0]), int(rt[0]), list(ms)

    def timing(self):
        """(instances on the wave-level kernel, instances on the workgroup-level
        kernel, [ms of the wave-level kernel, ms of the workgroup-level kernel
        with the copies out] in the last run)."""
        nw = np.zeros(1, np.int32)
        ng = np.zeros(1, np.int32)
        ms = (C.c_float * 2)()
        _check(lib().lassie_batch_plan_timing(self.h, _p(nw, _i32p), _p(ng, _i32p), ms))
        return int(nw[0]), int(ng[0]), list(ms)

    def progress(self):
        """Instances finished by the run in flight (safe from another thread)."""
        done = C.c_int64(0)
        _check(lib().lassie_batch_plan_progress(self.h, C.byref(done)))
        return int(done.value)

    def close(self):
        if getattr(self, "h", None):
            lib().lassie_batch_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
