"""Truncation plugins on the device (pq_set_keep_kernel): every branch is
formed on the device, the plugin decides on the host, and the store merge
applies the decisions with the pair rule's claims.  Against the oracle, which
calls the plugin's scalar keep child by child as the reference's emit gate
does (crates/core/src/engine/termsum.rs:129-148, 701-745): term keys and
coefficients bit-exact, per-gate live counts and branch counters equal."""

import ctypes
import math

import pytest

import propaq_amd as pa
from oracle import oracle as orc
from propaq_amd import runner
from tests.parity import build_truncator, compare, oracle_truncator, random_case

pytestmark = pytest.mark.gpu


def _plugin(tmp_path, name, **defs):
    path = build_truncator(tmp_path, name, **defs)
    nt = pa.NativeTruncator(path)
    lib = ctypes.CDLL(path)
    lib.tp_violations.restype = ctypes.c_uint64
    lib.tp_reset()
    return nt, lib


def _check_counters(d, o):
    for k in ("visited", "emitted", "declined", "hits"):
        assert d.counters[k] == o.counters[k], (k, d.counters[k], o.counters[k])
    assert math.isclose(d.counters["discarded_coeff_l1"], o.counters["discarded_coeff_l1"], rel_tol=1e-12,
                        abs_tol=1e-300)


@pytest.mark.parametrize("mode,kind,n_units,seed", [
    (0, 0, 10, 0), (0, 0, 40, 1), (0, 1, 16, 2),
    (1, 0, 12, 3), (1, 0, 70, 4), (1, 1, 20, 5),
    (2, 0, 30, 6), (2, 1, 24, 7),
])
def test_plugin_run_matches_oracle(tmp_path, mode, kind, n_units, seed):
    nt, lib = _plugin(tmp_path, f"m{mode}s{seed}", MODE=mode, THR="3e-3", WITH_BATCH=1)
    obs, layers, diag = random_case(800 + seed, kind, n_units, n_layers=8, gates_per_layer=10)
    print(f"mode {mode} kind {kind} n {n_units} seed {seed}", flush=True)
    d, o = compare(kind, n_units, obs, layers, damping=0.03, diag=diag, truncator=nt)
    _check_counters(d, o)
    assert o.counters["declined"] > 0
    assert lib.tp_violations() == 0


def test_scalar_only_plugin_and_term_budget(tmp_path):
    # no batch entry point: the host falls back to scalar calls; the term
    # budget keeps the cutoff lossless (no plugin call at emit) until it is met
    nt, lib = _plugin(tmp_path, "scalar", MODE=1, THR="1e-2")
    obs, layers, diag = random_case(820, 0, 10, n_layers=6, gates_per_layer=8)
    d, o = compare(0, 10, obs, layers, min_terms=40, diag=diag, truncator=nt)
    _check_counters(d, o)
    assert lib.tp_violations() == 0


def test_plugin_with_clifford_frame(tmp_path):
    # a plugin that does not read keys leaves Clifford deferral on; one that
    # does turns it off (crates/pauli/src/engine.rs:101-103)
    for mode in (0, 1):
        nt, _ = _plugin(tmp_path, f"cl{mode}", MODE=mode, THR="2e-3", WITH_BATCH=1)
        obs, layers, diag = random_case(830 + mode, 0, 16, n_layers=8, gates_per_layer=8, clifford_frac=0.5)
        d, o = compare(0, 16, obs, layers, diag=diag, truncator=nt)
        _check_counters(d, o)


def test_propagator_with_native_truncator(tmp_path):
    nt, lib = _plugin(tmp_path, "prop", MODE=2, THR="2e-3", WITH_BATCH=1)
    n = 14
    obs_terms, layers, diag = random_case(840, 0, n, n_layers=6, gates_per_layer=10)
    circ = pa.PauliCircuit([[pa.PauliRotation(pa.PauliString.from_key(g, n), a) for g, a in layer]
                            for layer in layers])
    obs = pa.PauliTermSum()
    for k, c in obs_terms:
        obs.add(pa.PauliString.from_key(k, n), c)
    prop = pa.PauliPropagator(noise=pa.UniformNoiseModel(0.02), truncation=[nt])
    res = prop.expectation_value(obs, circ, initial_state=diag)
    table = lambda li, nl: runner.uniform_table(0.02, n)  # noqa: E731
    o = orc.run_oracle(0, n, obs_terms, layers, noise_table=table, diag_state=diag, truncator=oracle_truncator(nt))
    assert res.n_terms == o.n_terms
    assert abs(res.expectation_value - o.expectation_value) <= 1e-12
    assert lib.tp_violations() == 0


def test_keep_kernel_refusals():
    from propaq_amd import _native

    eng = _native.DeviceEngine(orc.PAULI, 4, n_instances=2)
    with pytest.raises(RuntimeError, match="batched"):
        eng.set_keep_kernel(lambda *a: [1] * len(a[1]))
    eng.close()
    eng = _native.DeviceEngine(orc.PAULI, 4)
    eng.set_differentiable(True)
    eng.set_keep_kernel(lambda kind, keys, w, m, nu, li, nl: (m >= 0).astype("uint8"))
    with pytest.raises(RuntimeError, match="keep kernel"):
        eng.tape_begin()
    eng.close()
