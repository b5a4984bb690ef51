"""Structured event log of a device run (propagators with a Logger) against
the oracle driven gate by gate: per-gate live counts, new rows, declined
branches and discarded coefficient mass (the figures of
crates/pauli/src/engine.rs:165-193 and termsum.rs:611-626)."""

import math

import numpy as np
import pytest

import propaq_amd as pa
from oracle import oracle as orc
from propaq_amd import workloads

pytestmark = pytest.mark.gpu


def test_gate_log_matches_oracle_counters(tmp_path):
    n, steps = 10, 3
    rng = np.random.default_rng(2)
    layers = []
    for _ in range(steps):
        layers.append([pa.PauliRotation(pa.PauliString(0, (1 << q) | (1 << (q + 1)), n), float(rng.uniform(-3, 3)),
                                        qiskit_gate_idx=100 + q) for q in range(n - 1)])
        layers.append([pa.PauliRotation(pa.PauliString(1 << q, 0, n), float(rng.uniform(-3, 3))) for q in range(n)])
    circ = pa.PauliCircuit(layers)
    obs = pa.PauliTermSum()
    obs.add(pa.PauliString(0, 1 << 4, n), 1.0)
    log = pa.Logger(str(tmp_path / "run.jsonl"))
    prop = pa.PauliPropagator(truncation=pa.TruncationPolicy(5, 1e-3), logger=log)
    res = prop.expectation_value(obs, circ)
    p = pa.LogParser(log.filename)
    assert p.terms == res.n_terms
    # the oracle, gate by gate in Heisenberg order
    o = orc.OracleEngine(orc.PAULI, n)
    o.set_cutoff(5, 1e-3)
    o.add(workloads.z(4), 1.0)
    heis = [list(reversed(layer)) for layer in reversed(layers)]
    flat = [(li, r) for li, layer in enumerate(heis) for r in layer]
    assert len(p.truncation_events) == len(flat)
    prev = o.counters()
    for ev, g, (li, r) in zip(p.truncation_events, p.gate_events, flat):
        before = len(o)
        o.apply_rotation(r.generator.key, r.angle)
        c = o.counters()
        assert ev.layer_idx == li and g.qiskit_gate_idx == r.qiskit_gate_idx
        assert (ev.terms_before, ev.terms_after) == (before, len(o))
        assert ev.terms_gained == ev.terms_after - ev.terms_before
        assert ev.terms_discarded == c["declined"] - prev["declined"]
        # the log writes coefficients as {:.6e} (7 significant digits), as the reference does
        assert math.isclose(ev.discarded_coeff_l1, c["discarded_coeff_l1"] - prev["discarded_coeff_l1"],
                            rel_tol=1e-6, abs_tol=1e-15)
        assert math.isclose(ev.discarded_coeff_max, c["discarded_coeff_max"], rel_tol=1e-6, abs_tol=1e-15)
        assert ev.weight_cutoff == 5 and ev.coeff_cutoff == 1e-3 and ev.trigger == "emit"
        prev = c
    assert sum(p.terms_discarded) > 0 and sum(p.discarded_coeff_l1) > 0  # the case exercises the discard path
    (ph,) = p.engine_phases_events
    assert ph.terms == res.n_terms[-1] and ph.declined == sum(p.terms_discarded)
    # logging does not change the run
    plain = pa.PauliPropagator(truncation=pa.TruncationPolicy(5, 1e-3)).expectation_value(obs, circ)
    assert plain.expectation_value == res.expectation_value and plain.n_terms == res.n_terms
