"""Batched differentiation (pq_adjoint_seed_batched, pq_adjoint_backward_batched):
every instance's value and gradient against ``adjoint.run_adjoint`` on that
instance alone, whatever the checkpoint interval and tape budget: bit for bit
when the batch records on the chained kernels (the batched sweep reverses
each instance's rows in the blocks of a run of its own), values bit for bit
and gradients to rounding when it records on the sparse passes; and, without
truncation, against parameter shift of the oracle."""

import math

import numpy as np
import pytest

from oracle import oracle as orc
from propaq_amd import adjoint, batch, runner, workloads

pytestmark = pytest.mark.gpu


def hea_instances(n, depths, seed):
    rng = np.random.default_rng(seed)
    return [workloads.hea(n, d, rng) for d in depths]


def check_against_solo(kind, n, obs, inst, sparse=False, **kw):
    from propaq_amd import _native

    eng = _native.DeviceEngine(kind, n, n_instances=len(inst))
    r = batch.run_batch_gradient(kind, n, obs, inst, engine=eng, sparse=sparse, **kw)
    c = eng.counters()
    eng.close()
    if sparse:
        assert c["sparse_gates"] > 0 and c["chained_gates"] == 0, c
    else:
        assert c["sparse_gates"] == 0 and c["chained_gates"] > 0, c
    for b, layers in enumerate(inst):
        # the batched sweep reverses each instance in the blocks of a solo run on
        # the chained kernels, so that recording path is the bit-exact match
        s = adjoint.run_adjoint(kind, n, obs, layers, max_weight=kw.get("max_weight"),
                                min_coeff=kw.get("min_coeff"), checkpoint_every=kw.get("checkpoint_every", 4),
                                sparse=False)
        print(f"instance {b}: depth {len(layers)} <O> {s.expectation_value:.12f} |grad| {np.linalg.norm(s.gradient):.6e}",
              flush=True)
        assert r.expectation_values[b] == s.expectation_value, b
        if sparse:  # the list order of the sparse passes varies: partial sums agree to rounding
            scale = max(np.max(np.abs(s.gradient)), 1e-300)
            np.testing.assert_allclose(r.gradients[b], s.gradient, rtol=1e-11, atol=1e-14 * scale)
        else:
            assert np.array_equal(r.gradients[b], s.gradient), (b, np.max(np.abs(r.gradients[b] - s.gradient)))
    return r


@pytest.mark.parametrize("sparse", [False, True])
@pytest.mark.parametrize("n,ckpt", [(10, 4), (10, 1), (20, None)])
def test_batched_gradient_equals_solo_runs(n, ckpt, sparse):
    inst = hea_instances(n, [1, 3, 2, 3], seed=n)
    check_against_solo(orc.PAULI, n, workloads.tfim_observable(n), inst, sparse=sparse, max_weight=6,
                       min_coeff=1e-4, checkpoint_every=ckpt)


@pytest.mark.parametrize("sparse", [False, True])
def test_batched_gradient_with_gate_level_segments(sparse):
    # a tape budget far below one layer's tape cuts segments between gates
    n = 12
    inst = hea_instances(n, [2, 1, 3], seed=3)
    r = check_against_solo(orc.PAULI, n, workloads.tfim_observable(n), inst, sparse=sparse, max_weight=6,
                           min_coeff=1e-4, checkpoint_every=2, tape_budget=1 << 16)
    assert r.checkpoints > 2


def test_batched_sparse_gradient_of_wide_instances():
    # W = 2 keys (50 qubits and the instance bits), stores of thousands of rows
    n = 50
    inst = hea_instances(n, [2, 1, 3, 2, 1], seed=21)
    check_against_solo(orc.PAULI, n, workloads.tfim_observable(n), inst, sparse=True, max_weight=6,
                       min_coeff=1e-4, checkpoint_every=2)


def test_batched_gradient_matches_parameter_shift():
    # no truncation: the expectation is a + b cos + c sin in each angle, so
    # parameter shift with oracle forward runs is the exact derivative
    n = 6
    inst = hea_instances(n, [1, 2], seed=11)
    obs = workloads.tfim_observable(n)
    r = batch.run_batch_gradient(orc.PAULI, n, obs, inst)
    for b, layers in enumerate(inst):
        flat = [(g, a) for layer in layers for g, a in layer]
        for j in range(0, len(flat), 5):
            def value(d):
                out, pos = [], 0
                for layer in layers:
                    out.append([(g, a + d if pos + i == j else a) for i, (g, a) in enumerate(layer)])
                    pos += len(layer)
                return orc.run_oracle(orc.PAULI, n, obs, out).expectation_value
            ps = (value(math.pi / 2) - value(-math.pi / 2)) / 2.0
            assert abs(r.gradients[b][j] - ps) <= 1e-10, (b, j, r.gradients[b][j], ps)
        assert abs(r.expectation_values[b] - orc.run_oracle(orc.PAULI, n, obs, layers).expectation_value) <= 1e-10


def test_differentiable_values_agree_with_plain_runs():
    # the CZ ladders' Z(pi/2) rotations fold into the frame of a plain run; a
    # recorded run branches them with no truncation predicate, so no branch
    # below the coefficient cutoff is lost and the values agree
    n = 12
    obs = workloads.tfim_observable(n)
    inst = hea_instances(n, [2, 3], seed=9)
    for layers in inst:
        plain = runner.run_device(orc.PAULI, n, obs, layers, max_weight=6, min_coeff=1e-4)
        rec = adjoint.run_adjoint(orc.PAULI, n, obs, layers, max_weight=6, min_coeff=1e-4)
        assert abs(plain.expectation_value - rec.expectation_value) <= 1e-12
    plain = batch.run_batch(orc.PAULI, n, obs, inst, max_weight=6, min_coeff=1e-4)
    rec = batch.run_batch_gradient(orc.PAULI, n, obs, inst, max_weight=6, min_coeff=1e-4)
    assert np.max(np.abs(plain.expectation_values - rec.expectation_values)) <= 1e-12


def test_retire_is_refused_while_recording():
    from propaq_amd import _native

    eng = _native.DeviceEngine(orc.PAULI, 4, n_instances=2)
    eng.set_differentiable(True)
    eng.seed_batched(np.array([0, 1], dtype=np.uint32), np.array([[workloads.z(0)], [workloads.z(1)]],
                                                                    dtype=np.uint64), np.array([1.0, 1.0]))
    eng.tape_begin()
    with pytest.raises(RuntimeError, match="retired while recording"):
        eng.retire_instances(np.array([1, 0], dtype=np.uint8))
    eng.tape_end()
    eng.close()
