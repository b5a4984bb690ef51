"""Truncation plugins (NativeTruncator, EmitCutoff::term) on the oracle and the
host side.

The oracle's plugin path is pinned by what the reference's rules fix:
* a plugin that keeps iff the branch magnitude reaches THR takes the same
  decisions as a coefficient cutoff of THR with no weight bound, and with a
  plugin there is no precheck (EmitPrecheck::for_gate is None,
  crates/core/src/engine/termsum.rs:176-190), whose held-back value differs
  from the checked branch only by the unit phase, so both runs must agree bit
  for bit, counters included;
* a plugin that keeps everything is a lossless run;
* the plugin's own checks of its arguments (the weight of the raw key, words
  NULL exactly when it does not read keys, layer_index < n_layers) see no
  violation, and the layer count it sees is the circuit's.
"""

import ctypes

import numpy as np
import pytest

import propaq_amd as pa
from oracle import oracle as orc
from tests.parity import build_truncator, oracle_truncator, random_case


def _plugin(tmp_path, name, **defs):
    path = build_truncator(tmp_path, name, **defs)
    nt = pa.NativeTruncator(path)
    lib = ctypes.CDLL(path)
    for f, t in (("tp_calls", ctypes.c_uint64), ("tp_violations", ctypes.c_uint64), ("tp_max_layers", ctypes.c_uint32)):
        getattr(lib, f).restype = t
    lib.tp_reset()
    return nt, lib


@pytest.mark.parametrize("kind,n_units,seed", [(0, 10, 0), (0, 40, 1), (1, 12, 2), (0, 70, 3)])
def test_oracle_threshold_plugin_equals_coefficient_cutoff(tmp_path, kind, n_units, seed):
    nt, lib = _plugin(tmp_path, f"thr{seed}", MODE=0, THR="2e-3")
    obs, layers, diag = random_case(700 + seed, kind, n_units, n_layers=8, gates_per_layer=10)
    table = lambda li, nl: [0.97 ** w for w in range(n_units + 1)]  # noqa: E731
    a = orc.run_oracle(kind, n_units, obs, layers, noise_table=table, diag_state=diag, want_terms=True,
                       truncator=oracle_truncator(nt))
    b = orc.run_oracle(kind, n_units, obs, layers, min_coeff=2e-3, noise_table=table, diag_state=diag,
                       want_terms=True)
    assert a.n_terms == b.n_terms
    assert a.terms == b.terms
    assert a.counters == b.counters
    assert a.expectation_value == b.expectation_value
    assert b.counters["declined"] > 0 and lib.tp_calls() > 0 and lib.tp_violations() == 0


def test_oracle_keep_everything_plugin_is_lossless(tmp_path):
    nt, lib = _plugin(tmp_path, "all", MODE=1, THR="1e-300")
    obs, layers, diag = random_case(720, 0, 8, n_layers=5, gates_per_layer=6)
    a = orc.run_oracle(0, 8, obs, layers, diag_state=diag, want_terms=True, truncator=oracle_truncator(nt))
    b = orc.run_oracle(0, 8, obs, layers, diag_state=diag, want_terms=True, defer=False)
    assert a.terms == b.terms and a.n_terms == b.n_terms
    assert a.counters["declined"] == 0 and lib.tp_violations() == 0


@pytest.mark.parametrize("mode", [1, 2])
def test_oracle_plugin_arguments(tmp_path, mode):
    nt, lib = _plugin(tmp_path, f"args{mode}", MODE=mode, THR="5e-3")
    obs, layers, diag = random_case(730 + mode, 0, 30, n_layers=7, gates_per_layer=8)
    table = lambda li, nl: [0.98 ** w for w in range(31)]  # noqa: E731
    r = orc.run_oracle(0, 30, obs, layers, noise_table=table, diag_state=diag, truncator=oracle_truncator(nt))
    assert lib.tp_calls() > 100
    assert lib.tp_violations() == 0
    assert lib.tp_max_layers() == len(layers)
    assert r.counters["declined"] > 0


def test_oracle_term_budget_floor_switches_the_plugin_off(tmp_path):
    # below min_terms the cutoff is lossless (EmitCutoff::at_size): a plugin
    # that refuses most branches declines nothing until the floor is met
    # (seeds, admitted with the full cutoff, all reach its threshold)
    nt, _ = _plugin(tmp_path, "few", MODE=0, THR="0.15")
    obs, layers, diag = random_case(740, 0, 10, n_layers=6, gates_per_layer=8)
    r = orc.run_oracle(0, 10, obs, layers, min_terms=30, diag_state=diag, truncator=oracle_truncator(nt))
    lossless = orc.run_oracle(0, 10, obs, layers, diag_state=diag)
    first = next(i for i, n in enumerate(lossless.n_terms) if n >= 30)
    assert r.n_terms[:first + 1] == lossless.n_terms[:first + 1]
    assert max(r.n_terms) < max(lossless.n_terms)


def test_native_truncator_batch_and_scalar_decisions_agree(tmp_path):
    scalar, _ = _plugin(tmp_path, "s", MODE=1, THR="0.3")
    batched, _ = _plugin(tmp_path, "b", MODE=1, THR="0.3", WITH_BATCH=1)
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 1 << 62, size=(300, 2), dtype=np.uint64)
    w = rng.integers(0, 40, size=300).astype(np.uint32)
    m = rng.random(300)
    a = scalar.keep_flags(0, keys, w, m, 64, 1, 4)
    b = batched.keep_flags(0, keys, w, m, 64, 1, 4)
    assert np.array_equal(a, b) and 0 < a.sum() < 300
