"""Batched engine (pq_engine_create_batched ...): every instance of a batch
against a run of that instance alone on the device (keys, coefficients and
expectation values equal bit for bit) and against the oracle."""

import math

import numpy as np
import pytest

from oracle import oracle as orc
from propaq_amd import batch, runner, workloads

pytestmark = pytest.mark.gpu


def single(kind, n, obs, layers, diag, **kw):
    d = runner.run_device(kind, n, obs, layers, diag_state=diag, want_terms=True, **kw)
    return d.expectation_value, d.terms(), d.n_terms[-1] if d.n_terms else None


def check_batch(kind, n, observables, instances, diags, **kw):
    table = kw.pop("noise_table", None)
    engine = kw.pop("engine", None)
    run = batch.run_batch(kind, n, observables, instances, noise_table=table, diag_states=diags, want_terms=True,
                          engine=engine, **kw)
    B = len(instances)
    for b in range(B):
        obs = observables if isinstance(observables[0][0], int) else observables[b]
        d = diags if np.isscalar(diags) else diags[b]
        ev, terms, _ = single(kind, n, obs, instances[b], d,
                              noise_table=(lambda li, nl: table) if table is not None else None, **kw)
        print(f"instance {b}: depth {len(instances[b])} rows {len(terms)} <O> {ev}", flush=True)
        assert run.expectation_values[b] == ev, (b, run.expectation_values[b], ev)
        assert run.terms[b] == terms, b
        assert run.final_counts[b] == len(terms)
    return run


def hea_instances(n, depths, seed):
    rng = np.random.default_rng(seed)
    return [workloads.hea(n, d, rng) for d in depths]


@pytest.mark.parametrize("sparse", [True, False])
@pytest.mark.parametrize("n", [12, 32, 50, 130])
def test_hea_batch_matches_single_runs(n, sparse):
    # the batched store on the sparse passes (the default) and on the sorted
    # chained kernels: both give every instance's solo run bit for bit
    from propaq_amd import _native

    inst = hea_instances(n, [1, 3, 2, 3, 1, 2, 4], seed=n)
    eng = _native.DeviceEngine(orc.PAULI, n, n_instances=len(inst))
    eng.set_sparse(sparse)
    run = check_batch(orc.PAULI, n, workloads.tfim_observable(n), inst, 0, max_weight=6, min_coeff=1e-4,
                      engine=eng)
    c = eng.counters()
    eng.close()
    assert run.steps == 16
    if sparse:
        assert c["sparse_gates"] > 0 and c["chained_gates"] == 0, c
    else:
        assert c["sparse_gates"] == 0 and c["chained_gates"] > 0, c


@pytest.mark.parametrize("n", [12, 50])
def test_batch_readout_of_unsorted_store(n):
    # without term export the finished instances are read out of the store
    # the sparse passes left unsorted (pq_readout_retire sorts only theirs):
    # values and row counts equal the solo runs'
    inst = hea_instances(n, [1, 3, 2, 3, 1, 2, 4, 2], seed=100 + n)
    obs = workloads.tfim_observable(n)
    run = batch.run_batch(orc.PAULI, n, obs, inst, max_weight=6, min_coeff=1e-4)
    for b, layers in enumerate(inst):
        d = runner.run_device(orc.PAULI, n, obs, layers, max_weight=6, min_coeff=1e-4)
        assert run.expectation_values[b] == d.expectation_value, b
        assert run.final_counts[b] == (d.n_terms[-1] if d.n_terms else len(obs)), b


def test_per_instance_observables_states_noise_and_phase_only():
    n = 10
    rng = np.random.default_rng(3)
    inst = hea_instances(n, [2, 1, 3, 2], seed=5)
    # phase-only angles (0, pi) in some instances only
    for b, layers in enumerate(inst):
        for layer in layers[:4]:
            for j in range(0, len(layer), 3):
                g, a = layer[j]
                if abs(abs(a) - math.pi / 2) > 1e-6:
                    layer[j] = (g, [0.0, math.pi, a, a][b])
    obs = [[(workloads.z(int(rng.integers(n))) | workloads.x(int(rng.integers(n))), float(rng.uniform(0.5, 1.5))),
            (workloads.z(b), 1.0)] for b in range(4)]
    diags = [int(rng.integers(0, 1 << n)) for _ in range(4)]
    table = runner.uniform_table(0.02, n)
    check_batch(orc.PAULI, n, obs, inst, diags, max_weight=5, min_coeff=1e-3, noise_table=table)


def majorana_instances(n_sites, depths, seed):
    rng = np.random.default_rng(seed)
    out = []
    for d in depths:
        layers = []
        for _ in range(d):
            for start in (0, 1):
                layer = []
                for p in range(start, n_sites - 1, 2):
                    # XX+YY Givens pair as two Majorana rotations on modes 2p.., and a density term
                    th = float(rng.uniform(-math.pi, math.pi))
                    layer.append(((1 << (2 * p + 1)) | (1 << (2 * p + 2)), th))
                    layer.append(((1 << (2 * p)) | (1 << (2 * p + 3)), -th))
                layers.append(layer)
            layers.append([((1 << (2 * p)) | (1 << (2 * p + 1)), float(rng.uniform(-math.pi, math.pi)))
                           for p in range(n_sites)])
        out.append(layers)
    return out


def test_majorana_batch_matches_single_runs():
    n = 8
    inst = majorana_instances(n, [1, 2, 3, 2], seed=7)
    obs = [((1 << (2 * p)) | (1 << (2 * p + 1)), 0.5) for p in range(4)] + [(0, 2.0)]
    check_batch(orc.MAJORANA, n, obs, inst, (1 << 4) - 1, max_weight=4, min_coeff=1e-4)


def test_batch_against_oracle():
    n = 12
    inst = hea_instances(n, [2, 1, 3], seed=1)
    obs = workloads.tfim_observable(n)
    run = batch.run_batch(orc.PAULI, n, obs, inst, max_weight=5, min_coeff=1e-4)
    for b, layers in enumerate(inst):
        o = orc.run_oracle(orc.PAULI, n, obs, layers, max_weight=5, min_coeff=1e-4)
        assert abs(run.expectation_values[b] - o.expectation_value) <= 1e-10


def test_mismatched_structure_is_refused():
    n = 6
    inst = hea_instances(n, [1, 1], seed=2)
    inst[1][0][0] = (workloads.x(0), 0.3)
    with pytest.raises(ValueError, match="different generators"):
        batch.run_batch(orc.PAULI, n, workloads.tfim_observable(n), inst)
