"""Device adjoint gradients (pq_tape_* / pq_adjoint_*) against independent
derivatives computed with the oracle.

* Without truncation the expectation is, in each angle, a + b cos + c sin
  (every generator squares to the identity), so parameter shift,
  (f(theta + pi/2) - f(theta - pi/2)) / 2, is the exact derivative.  It is
  evaluated with oracle forward runs (the oracle itself is pinned against
  dense matrices), angle by angle, including Clifford angles (+-pi/2, 3pi/2)
  and phase-only angles (0, +-pi), where differentiable mode must still carry
  the branch.
* With truncation the gradient is that of the computed function with its
  decisions held fixed; central finite differences of the oracle (h = 1e-6,
  cross-checked at h / 4 so that an angle whose perturbation flips a
  truncation decision is recognised and left out) pin it.
* Recomputation from checkpoints is exact: the gradient does not depend on
  the checkpoint interval, bit for bit when the gates record on the sorted
  chained kernels and to rounding when they record on the sparse passes
  (whose gradient sums follow the order the lists were filled in); the two
  recording paths agree to rounding.
"""

import math

import numpy as np
import pytest

from oracle import oracle as orc
from propaq_amd import adjoint, runner, workloads
from tests.parity import random_case

pytestmark = pytest.mark.gpu


def oracle_value(kind, n, obs, layers, diag, **kw):
    return orc.run_oracle(kind, n, obs, layers, diag_state=diag, **kw).expectation_value


def shifted(layers, flat_index, delta):
    out, pos = [], 0
    for layer in layers:
        row = []
        for g, a in layer:
            row.append((g, a + delta if pos == flat_index else a))
            pos += 1
        out.append(row)
    return out


def parameter_shift(kind, n, obs, layers, diag, indices, **kw):
    return np.array([(oracle_value(kind, n, obs, shifted(layers, j, math.pi / 2), diag, **kw)
                      - oracle_value(kind, n, obs, shifted(layers, j, -math.pi / 2), diag, **kw)) / 2.0
                     for j in indices])


def table_fn(damping, n):
    return (lambda li, nl: runner.uniform_table(damping, n)) if damping is not None else None


@pytest.mark.parametrize("kind,n", [(orc.PAULI, 8), (orc.MAJORANA, 5)])
@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("damping", [None, 0.05])
def test_lossless_gradient_matches_parameter_shift(kind, n, seed, damping):
    obs, layers, diag = random_case(100 + seed, kind, n, n_layers=5, gates_per_layer=5,
                                    clifford_frac=0.2, phase_frac=0.15)
    print(f"case kind={kind} n={n} seed={seed} damping={damping}", flush=True)
    table = table_fn(damping, n)
    run = adjoint.run_adjoint(kind, n, obs, layers, noise_table=table, diag_state=diag, checkpoint_every=2)
    ref = oracle_value(kind, n, obs, layers, diag, noise_table=table, defer=False)
    assert abs(run.expectation_value - ref) <= 1e-10
    n_rot = sum(len(l) for l in layers)
    ps = parameter_shift(kind, n, obs, layers, diag, range(n_rot), noise_table=table, defer=False)
    np.testing.assert_allclose(run.gradient, ps, rtol=0, atol=1e-10)


@pytest.mark.parametrize("sparse", [False, True])
@pytest.mark.parametrize("kind,n", [(orc.PAULI, 10), (orc.MAJORANA, 6)])
def test_gradient_does_not_depend_on_checkpoint_interval(kind, n, sparse):
    obs, layers, diag = random_case(7, kind, n, n_layers=7, gates_per_layer=6)
    kw = dict(max_weight=4, min_coeff=1e-3, noise_table=table_fn(0.02, n), diag_state=diag, sparse=sparse)
    runs = [adjoint.run_adjoint(kind, n, obs, layers, checkpoint_every=k, **kw) for k in (0, 1, 2, 3, 7)]
    # tiny tape budgets cut segments between single gates
    runs += [adjoint.run_adjoint(kind, n, obs, layers, checkpoint_every=0, tape_budget=b, **kw)
             for b in (1 << 12, 1 << 8)]
    scale = np.max(np.abs(runs[0].gradient))
    for r in runs[1:]:
        assert r.expectation_value == runs[0].expectation_value
        if sparse:
            np.testing.assert_allclose(r.gradient, runs[0].gradient, rtol=1e-12, atol=1e-15 * scale)
        else:
            assert np.array_equal(r.gradient, runs[0].gradient)
    assert runs[1].checkpoints == 6 and runs[0].checkpoints == 0
    assert runs[-1].checkpoints > runs[-2].checkpoints
    assert runs[-1].checkpoints >= 30  # about one segment per gate


@pytest.mark.parametrize("kind,n,seed", [(orc.PAULI, 12, 0), (orc.PAULI, 40, 1), (orc.MAJORANA, 8, 2)])
def test_sparse_and_chained_recording_agree(kind, n, seed):
    """The same session recorded on the sparse passes and on the chained
    kernels: equal values, gradients equal to rounding, and the sparse run
    really recorded sparse gates (with their sorts at the tape ends)."""
    from propaq_amd import _native

    obs, layers, diag = random_case(500 + seed, kind, n, n_layers=8, gates_per_layer=10, clifford_frac=0.1)
    kw = dict(max_weight=6 if kind == orc.PAULI else 4, min_coeff=1e-4, noise_table=table_fn(0.01, n),
              diag_state=diag, checkpoint_every=3)
    eng = _native.DeviceEngine(kind, n)
    sp = adjoint.run_adjoint(kind, n, obs, layers, engine=eng, sparse=True, **kw)
    c = eng.counters()
    eng.close()
    assert c["sparse_gates"] > 0 and c["chained_gates"] == 0, c
    ch = adjoint.run_adjoint(kind, n, obs, layers, sparse=False, **kw)
    assert sp.expectation_value == ch.expectation_value
    scale = max(np.max(np.abs(ch.gradient)), 1e-300)
    np.testing.assert_allclose(sp.gradient, ch.gradient, rtol=1e-11, atol=1e-14 * scale)


def test_sparse_recording_falls_back_to_chained_mid_tape():
    """Recording on the sparse passes until the store outgrows a small sparse
    budget, then on the chained kernels, inside one tape: the sort the
    chained kernels need is recorded before the gate that needs it, so the
    gradients agree with a chained-only recording to rounding."""
    from propaq_amd import _native

    w = workloads.config(4)
    layers = w.layers[:6]
    kw = dict(max_weight=w.max_weight, min_coeff=w.min_coeff, diag_state=w.diag_state, checkpoint_every=None)
    eng = _native.DeviceEngine(w.kind, w.n_units)
    eng.set_sparse_budget(1 << 20)  # stores of more than a few thousand rows go chained
    sp = adjoint.run_adjoint(w.kind, w.n_units, w.observable, layers, engine=eng, sparse=True, **kw)
    c = eng.counters()
    eng.close()
    print(c, flush=True)
    assert c["sparse_gates"] > 0 and c["chained_gates"] > 0 and c["sparse_fallbacks"] > 0, c
    ch = adjoint.run_adjoint(w.kind, w.n_units, w.observable, layers, sparse=False, **kw)
    assert sp.expectation_value == ch.expectation_value
    scale = max(np.max(np.abs(ch.gradient)), 1e-300)
    np.testing.assert_allclose(sp.gradient, ch.gradient, rtol=1e-11, atol=1e-14 * scale)


def finite_difference_check(kind, n, obs, layers, diag, grad, indices, h=1e-6, **kw):
    """Central differences of the oracle; angles whose differences at h and
    h / 4 disagree (a truncation decision flipped) are skipped."""
    checked = 0
    for j in indices:
        fd = []
        for step in (h, h / 4):
            fp = oracle_value(kind, n, obs, shifted(layers, j, step), diag, **kw)
            fm = oracle_value(kind, n, obs, shifted(layers, j, -step), diag, **kw)
            fd.append((fp - fm) / (2 * step))
        if abs(fd[0] - fd[1]) > 1e-6 * max(1.0, abs(fd[0])):
            continue
        checked += 1
        assert abs(grad[j] - fd[1]) <= 2e-6 * max(1.0, abs(fd[1])), (j, grad[j], fd)
    return checked


@pytest.mark.parametrize("kind,n", [(orc.PAULI, 10), (orc.MAJORANA, 6)])
@pytest.mark.parametrize("seed", range(3))
def test_truncated_gradient_matches_finite_differences(kind, n, seed):
    obs, layers, diag = random_case(200 + seed, kind, n, n_layers=6, gates_per_layer=6,
                                    clifford_frac=0.0, phase_frac=0.0)
    print(f"case kind={kind} n={n} seed={seed}", flush=True)
    kw = dict(max_weight=5 if kind == orc.PAULI else 3, min_coeff=2e-3, noise_table=table_fn(0.03, n))
    run = adjoint.run_adjoint(kind, n, obs, layers, diag_state=diag, checkpoint_every=2, **kw)
    ref = oracle_value(kind, n, obs, layers, diag, defer=False, **kw)
    assert abs(run.expectation_value - ref) <= 1e-10
    n_rot = sum(len(l) for l in layers)
    checked = finite_difference_check(kind, n, obs, layers, diag, run.gradient, range(n_rot), **kw)
    assert checked >= 0.8 * n_rot


@pytest.mark.parametrize("n", [40, 100, 200])
def test_wide_keys_lossless_gradient(n):
    obs, layers, diag = random_case(9, orc.PAULI, n, n_layers=4, gates_per_layer=10)
    run = adjoint.run_adjoint(orc.PAULI, n, obs, layers, diag_state=diag, checkpoint_every=1)
    rng = np.random.default_rng(n)
    idx = rng.choice(sum(len(l) for l in layers), size=12, replace=False)
    ps = parameter_shift(orc.PAULI, n, obs, layers, diag, idx, defer=False)
    np.testing.assert_allclose(run.gradient[idx], ps, rtol=0, atol=1e-10)


def test_differentiation_workload_prefix():
    """The 10 x 10 kicked-Ising differentiation workload (independent random
    angles, w <= 10, |c| >= 1e-4), first 2 Trotter steps: value against the
    oracle and 16 seeded angles against central differences."""
    w = workloads.config(4)
    layers = w.layers[:4]
    run = adjoint.run_adjoint(w.kind, w.n_units, w.observable, layers, max_weight=w.max_weight,
                              min_coeff=w.min_coeff, diag_state=w.diag_state, checkpoint_every=1)
    kw = dict(max_weight=w.max_weight, min_coeff=w.min_coeff)
    ref = oracle_value(w.kind, w.n_units, w.observable, layers, w.diag_state, defer=False, **kw)
    assert abs(run.expectation_value - ref) <= 1e-10
    n_rot = sum(len(l) for l in layers)
    idx = np.random.default_rng(4).choice(n_rot, size=16, replace=False)
    checked = finite_difference_check(w.kind, w.n_units, w.observable, layers, w.diag_state, run.gradient, idx, **kw)
    assert checked >= 12


def test_propagator_gradient_and_torch_function():
    import torch

    import propaq_amd as pa

    n = 6
    rng = np.random.default_rng(11)
    layers = []
    for _ in range(3):
        layers.append([pa.PauliRotation(pa.PauliString(0, (1 << q) | (1 << (q + 1)), n), float(rng.uniform(-3, 3)))
                       for q in range(n - 1)])
        layers.append([pa.PauliRotation(pa.PauliString(1 << q, 0, n), float(rng.uniform(-3, 3))) for q in range(n)])
    circ = pa.PauliCircuit(layers)
    obs = pa.PauliTermSum()
    obs.add(pa.PauliString(0, 1 << 2, n), 1.0)
    obs.add(pa.PauliString(0b11, 0, n), 0.5)
    prop = pa.PauliPropagator()
    res, grad = prop.expectation_value_and_gradient(obs, circ, checkpoint_every=1)
    plain = prop.expectation_value(obs, circ)
    assert abs(res.expectation_value - plain.expectation_value) <= 1e-12
    theta = torch.tensor([r.angle for layer in circ.layers for r in layer], dtype=torch.float64, requires_grad=True)
    val = pa.expectation(prop, obs, circ, theta)
    val.backward()
    assert abs(val.item() - res.expectation_value) <= 1e-12
    np.testing.assert_allclose(theta.grad.numpy(), grad, rtol=0, atol=1e-12)
    # the same derivative by parameter shift through the propagator
    j = 4
    angles = theta.detach().numpy().copy()
    vals = []
    for d in (math.pi / 2, -math.pi / 2):
        a = angles.copy()
        a[j] += d
        vals.append(pa.expectation(prop, obs, circ, torch.tensor(a)).item())
    assert abs((vals[0] - vals[1]) / 2 - grad[j]) <= 1e-10
