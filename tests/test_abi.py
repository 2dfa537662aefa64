"""The C ABI (include/lassie.h) on the CPU: the library loads, exports every
declared symbol, builds the reference's CSR, reports the reference's option
defaults, and fails loudly (no fallback) when no HIP device is visible."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from lassie_amd import _lib, synth

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "lassie.h")


def test_every_declared_symbol_is_exported():
    text = open(HEADER).read()
    declared = set(re.findall(r"\b(lassie_[a-z_]+)\s*\(", text))
    assert declared == set(_lib.EXPORTS)
    lib = C.CDLL(_lib.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name
    assert _lib.lib().lassie_abi_version() == 2


def test_option_defaults_are_the_references():
    o = _lib.default_opts()
    # Simulation.cuh:271-354
    assert (o.rkf_step, o.be_step, o.newton_iter, o.newton_tol, o.stiffness_tol) == \
        (1e-3, 1e-2, 10000, 1e-6, 1e-6)
    assert o.rtol == 0.0 and o.bdf_regrow == 0 and o.max_rkf_retries == 0
    assert o.precision == _lib.PREC_FP64 and o.max_newton == 0


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_csr_from_dense_matches_generator(seed):
    m = synth.random_model(17, 23, seed)
    got = _lib.csr_from_dense(m.L, m.Rm)
    for a, b in zip(got, (m.lptr, m.lsp, m.lexp, m.hptr, m.hrx, m.hcoef)):
        np.testing.assert_array_equal(a, b)


def test_csr_from_dense_entry_order_and_net_zero():
    # A + B -> A + C: A is a catalyst, so H has no entry for it (count_nonzero)
    L = np.array([[1, 1, 0], [0, 0, 2]], np.int16)
    Rm = np.array([[1, 0, 1], [1, 0, 0]], np.int16)
    lptr, lsp, lexp, hptr, hrx, hcoef = _lib.csr_from_dense(L, Rm)
    assert list(lptr) == [0, 2, 3] and list(lsp) == [0, 1, 2] and list(lexp) == [1, 1, 2]
    assert list(hptr) == [0, 1, 2, 4]
    assert list(hrx) == [1, 0, 0, 1] and list(hcoef) == [1, -1, 1, -2]


def test_invalid_arguments_are_rejected():
    L = _lib.lib()
    h = C.c_void_p()
    assert L.lassie_model_create(None, C.byref(h)) == _lib.LASSIE_ERR_INVALID
    assert b"NULL" in L.lassie_last_error()
    assert L.lassie_solve(None, None, None, 0, None, 0, None, None, None, 0, None) == _lib.LASSIE_ERR_INVALID


def test_no_device_fails_loudly():
    if _lib.device_count() > 0:
        pytest.skip("a HIP device is visible")
    with pytest.raises(_lib.LassieError) as e:
        _lib.DeviceModel(synth.small_nonstiff(0))
    assert e.value.code == _lib.LASSIE_ERR_HIP
