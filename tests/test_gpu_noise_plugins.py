"""Noise models through a device propagation, against the oracle (row f2):
the three evaluation strategies of crates/core/src/engine/noise_resolver.rs:
28-37 and their call pattern (tests/noise/test_native_plugins.py in the
reference):

* a native model that declares nothing is tabulated once per run (n_units + 1
  calls, never handed a key) and the table is uploaded;
* one that declares LAYER is re-tabulated at every layer start;
* one that declares KEY is called per term with the key words (factors go to
  the device through pq_scale_by_factors), and costs Clifford deferral;
* Python models: damping_factor (a table) and damping_factor_term (per term).

The oracle runs the same circuit with the same factors (the test asks the
plugin or the Python model itself for them, on the oracle's rows); terms must
agree bit for bit, live counts exactly, expectation values within 1e-10."""

import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

import propaq_amd as pa
from oracle import oracle as orc
from tests.parity import random_case

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _plugin(tmp_path, depends, batch=0):
    so = str(tmp_path / f"noise_{depends}_{batch}.so")
    subprocess.run(["gcc", "-shared", "-fPIC", "-O1", f"-DDEPENDS={depends}", f"-DWITH_BATCH={batch}", "-o", so,
                    os.path.join(HERE, "noise_plugin.c"), "-lm"], check=True)
    lib = ctypes.CDLL(so)
    for f in ("np_calls", "np_saw_words", "np_violations"):
        getattr(lib, f).restype = ctypes.c_uint64
    lib.np_max_layer.restype = ctypes.c_uint32
    return pa.NativeNoiseModel(so), lib


def _objects(kind, n, obs_terms, layers):
    if kind == 0:
        circ = pa.PauliCircuit([[pa.PauliRotation(pa.PauliString.from_key(g, n), a) for g, a in layer]
                                for layer in layers])
        obs = pa.PauliTermSum()
        for k, c in obs_terms:
            obs.add(pa.PauliString.from_key(k, n), c)
        return pa.PauliPropagator, obs, circ
    circ = pa.MajoranaCircuit([[pa.MajoranaRotation(pa.MajoranaMonomial.from_key(g, 2 * n), a) for g, a in layer]
                               for layer in layers])
    obs = pa.MajoranaTermSum()
    for k, c in obs_terms:
        obs.add(pa.MajoranaMonomial.from_key(k, 2 * n), c)
    return pa.MajoranaPropagator, obs, circ


def _raw_rows(eng):
    keys, _ = eng.export(frame=False)
    w = np.array([orc.weight(eng.kind, orc.from_words(k), eng.n_units, eng.W) for k in keys], dtype=np.uint32)
    return keys, w


CUT = dict(max_weight=5, min_coeff=2e-3)


def _device(kind, n, obs_terms, layers, noise, diag):
    prop_cls, obs, circ = _objects(kind, n, obs_terms, layers)
    prop = prop_cls(noise=noise, truncation=pa.TruncationPolicy(weight_cutoff=CUT["max_weight"],
                                                                 coeff_cutoff=CUT["min_coeff"]))
    evolved = prop.propagate(obs, circ)
    res = prop.expectation_value(obs, circ, initial_state=diag)
    return evolved, res


def _check(evolved, res, o):
    assert res.n_terms == o.n_terms, "per-gate live counts differ"
    assert evolved._terms == o.terms, "terms differ"
    assert abs(res.expectation_value - o.expectation_value) <= 1e-10


CASES = [(0, 12, 900), (0, 40, 901), (1, 16, 902)]


@pytest.mark.parametrize("kind,n,seed", CASES)
@pytest.mark.parametrize("batch", [0, 1])
def test_weight_only_native_model_is_tabulated_once(tmp_path, kind, n, seed, batch):
    model, lib = _plugin(tmp_path, 0, batch)
    obs, layers, diag = random_case(seed, kind, n, n_layers=6, gates_per_layer=10)
    table = model.table(kind, n)
    o = orc.run_oracle(kind, n, obs, layers, noise_table=lambda li, nl: table, diag_state=diag, want_terms=True,
                       **CUT)
    lib.np_reset()
    evolved, res = _device(kind, n, obs, layers, model, diag)
    _check(evolved, res, o)
    assert lib.np_calls() == 2 * (n + 1)  # once per run: propagate, then expectation_value
    assert lib.np_saw_words() == 0 and lib.np_violations() == 0


@pytest.mark.parametrize("kind,n,seed", CASES)
def test_layer_aware_native_model_is_retabulated_per_layer(tmp_path, kind, n, seed):
    model, lib = _plugin(tmp_path, 2)
    obs, layers, diag = random_case(seed, kind, n, n_layers=6, gates_per_layer=10)
    L = len(layers)
    tables = [model.table(kind, n, li, L) for li in range(L)]
    o = orc.run_oracle(kind, n, obs, layers, noise_table=lambda li, nl: tables[li], diag_state=diag,
                       want_terms=True, **CUT)
    lib.np_reset()
    evolved, res = _device(kind, n, obs, layers, model, diag)
    _check(evolved, res, o)
    assert lib.np_calls() == 2 * L * (n + 1)
    assert lib.np_max_layer() == L - 1
    assert lib.np_saw_words() == 0 and lib.np_violations() == 0


@pytest.mark.parametrize("kind,n,seed", CASES)
@pytest.mark.parametrize("batch", [0, 1])
def test_key_aware_native_model_is_called_per_term(tmp_path, kind, n, seed, batch):
    model, lib = _plugin(tmp_path, 3, batch)  # KEY and LAYER
    obs, layers, diag = random_case(seed, kind, n, n_layers=6, gates_per_layer=10)
    rows_at_layer_start = [0]

    def factors(eng, li, nl):
        rows_at_layer_start[0] += len(eng)
        keys, w = _raw_rows(eng)
        return model.factors(kind, keys, w, n, li, nl)
    o = orc.run_oracle(kind, n, obs, layers, noise_factors=factors, diag_state=diag, defer=False,
                       want_terms=True, **CUT)
    lib.np_reset()
    evolved, res = _device(kind, n, obs, layers, model, diag)
    _check(evolved, res, o)
    # one call per live term at every layer start, in each of the two runs
    calls = lib.np_calls()
    assert calls == 2 * rows_at_layer_start[0] > 0
    assert lib.np_saw_words() == calls and lib.np_violations() == 0


class _Weighted(pa.GateNoiseModel):
    def damping_factor(self, term_weight, active_modes=0):
        return math.exp(-0.04 * term_weight)


class _PerTerm(pa.GateNoiseModel):
    def damping_factor_term(self, kind, words, n_units, weight):
        return math.exp(-0.03 * weight) * (0.9 if words[0] & 0b11 else 1.0)


@pytest.mark.parametrize("kind,n,seed", CASES)
def test_python_weight_model(kind, n, seed):
    model = _Weighted()
    obs, layers, diag = random_case(seed, kind, n, n_layers=6, gates_per_layer=10)
    table = [model.damping_factor(w, 0) for w in range(n + 1)]
    o = orc.run_oracle(kind, n, obs, layers, noise_table=lambda li, nl: table, diag_state=diag, want_terms=True,
                       **CUT)
    _check(*_device(kind, n, obs, layers, model, diag), o)


@pytest.mark.parametrize("kind,n,seed", CASES)
def test_python_term_model(kind, n, seed):
    model = _PerTerm()
    obs, layers, diag = random_case(seed, kind, n, n_layers=6, gates_per_layer=10)

    def factors(eng, li, nl):
        keys, w = _raw_rows(eng)
        return np.array([model.damping_factor_term(kind, [int(v) for v in k], n, int(x)) for k, x in zip(keys, w)])
    o = orc.run_oracle(kind, n, obs, layers, noise_factors=factors, diag_state=diag, defer=False, want_terms=True,
                       **CUT)
    _check(*_device(kind, n, obs, layers, model, diag), o)


def _terms_after_clifford(noise):
    n = 4
    obs = pa.PauliTermSum({pa.PauliString(0, 0b0001, n): 1.0})
    circ = pa.PauliCircuit([pa.PauliRotation(pa.PauliString(0b0001, 0, n), math.pi / 2)])
    return len(pa.PauliPropagator(noise=noise).propagate(obs, circ))


def test_a_key_dependency_costs_clifford_deferral(tmp_path):
    """A key-reading model sees stored keys, so the Clifford branches instead
    of folding into the frame and its source row stays with cos(pi/2)."""
    assert _terms_after_clifford(_plugin(tmp_path, 0)[0]) == 1
    assert _terms_after_clifford(_plugin(tmp_path, 2)[0]) == 1
    assert _terms_after_clifford(_plugin(tmp_path, 1)[0]) == 2
    assert _terms_after_clifford(_PerTerm()) == 2
    assert _terms_after_clifford(_Weighted()) == 1
