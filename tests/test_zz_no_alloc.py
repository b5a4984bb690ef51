"""The allocation-free expectation entry point (pq_expectation_noalloc),
called through the C ABI, against pq_expectation on the same engine: bit for
bit on two calls in a row, the second on a smaller store, with a deferred
Clifford frame in play, and within 1e-10 of the oracle."""

import ctypes

import numpy as np
import pytest

from oracle import oracle as orc
from propaq_amd import _native, runner
from tests.parity import random_case

pytestmark = pytest.mark.gpu


def _noalloc(eng, diag: int) -> float:
    nd = max(1, (eng.n_units + 63) // 64)
    d = _native.int_to_words(int(diag), nd)
    out = ctypes.c_double(0.0)
    rc = _native.load().pq_expectation_noalloc(eng._h, d.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), nd,
                                                ctypes.byref(out))
    assert rc == 0, _native.load().pq_last_error()
    return float(out.value)


def _run(eng, obs, layers, diag, **cut):
    d = runner.run_device(0, eng.n_units, obs, layers, diag_state=diag, engine=eng, **cut)
    o = orc.run_oracle(0, eng.n_units, obs, layers, diag_state=diag, **cut)
    return d, o


@pytest.mark.parametrize("n_units", [12, 40, 100])
def test_noalloc_expectation_matches_the_allocating_entry_point(n_units):
    eng = _native.DeviceEngine(0, n_units)
    # first call: a larger store with deferred Cliffords in the frame
    obs, layers, diag = random_case(900 + n_units, 0, n_units, n_layers=8, gates_per_layer=12, clifford_frac=0.4)
    d, o = _run(eng, obs, layers, diag, min_coeff=1e-4)
    assert not eng.frame_is_identity()
    n_first = len(eng)
    v = _noalloc(eng, diag)
    assert v == eng.expectation(diag) == d.expectation_value
    assert abs(v - o.expectation_value) <= 1e-10
    # second call, same engine, a smaller store
    obs2, layers2, diag2 = random_case(950 + n_units, 0, n_units, n_layers=2, gates_per_layer=5, clifford_frac=0.4)
    d2, o2 = _run(eng, obs2, layers2, diag2, min_coeff=1e-2)
    assert len(eng) < n_first
    v2 = _noalloc(eng, diag2)
    assert v2 == eng.expectation(diag2) == d2.expectation_value
    assert abs(v2 - o2.expectation_value) <= 1e-10


def test_noalloc_expectation_of_an_empty_store():
    eng = _native.DeviceEngine(0, 8)
    eng.seed(np.zeros((0, 1), dtype=np.uint64), np.zeros(0))
    assert _noalloc(eng, 0) == 0.0
