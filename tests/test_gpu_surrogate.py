"""Device parity of the surrogate model (pq_surrogate_compile, pq_model_*):
the compiled tape evaluated by the gfx950 kernels equals the oracle
(oracle/surrogate.py) bit for bit, and the Python API behaves as the
reference's (tests/surrogate/*.py).
"""

from __future__ import annotations

import math
import os
import random

import numpy as np
import pytest

from oracle import surrogate as S
from oracle import oracle as O
from surrogate_cases import CONFIGS, random_case, random_params, wide_case

pytestmark = pytest.mark.gpu


def _device_model(kind, n, obs, layers, cfg, state, n_params, compact_min=None):
    from propaq_amd import _native
    from propaq_amd.runner import keys_to_array

    b = _native.SurrogateBuilder(kind, n)
    if compact_min is not None:
        b.set_compaction(compact_min)
    b.set_truncation(frequency=cfg.frequency, coefficient=cfg.coefficient, weight=cfg.weight,
                     min_terms=cfg.min_terms, simplify=cfg.simplify)
    b.seed(keys_to_array([k for k, _ in obs], b.W), np.array([c for _, c in obs]))
    gates = [g for layer in reversed(layers) for g in reversed(layer)]
    if gates:
        gens = keys_to_array([g for g, _, _ in gates], b.W)
        pidx = np.array([-1 if p is None else p for _, p, _ in gates], dtype=np.int64)
        ang = np.array([0.0 if a is None else a for _, _, a in gates])
        b.apply(gens, pidx, ang)
    m = b.compile(state, n_params)
    b.close()
    return m


@pytest.mark.parametrize("seed", range(14))
def test_device_model_matches_oracle(seed):
    kind = seed % 2
    cfg = CONFIGS[seed % len(CONFIGS)]
    n = 4 + seed % 3
    state = (seed * 5) % (1 << n)
    obs, layers = random_case(900 + seed, kind, n=n, n_gates=14)
    print(f"surrogate parity seed={seed} kind={kind} n={n}", flush=True)
    ref = S.build(kind, n, obs, layers, cfg, initial_state=state, n_params=5)
    dm = _device_model(kind, n, obs, layers, cfg, state, 5, compact_min=16 if seed % 3 == 0 else None)
    assert dm.n_terms == ref.n_terms
    assert dm.n_monomials == ref.n_monomials
    ps = random_params(seed, 7, 5)
    want = [ref.evaluate(p) for p in ps]
    assert list(dm.evaluate(np.array(ps))) == want
    assert dm.evaluate(np.array(ps[2:3]))[0] == want[2]


@pytest.mark.parametrize("kind", [S.PAULI, S.MAJORANA])
@pytest.mark.parametrize("n", [40, 100, 200])
def test_device_model_wide_keys(kind, n):
    obs, layers = wide_case(kind, n, n + kind)
    state = 1 | (1 << (n // 2)) | (1 << (n - 1))
    print(f"surrogate wide keys kind={kind} n={n}", flush=True)
    ref = S.build(kind, n, obs, layers, S.Config(frequency=6), initial_state=state, n_params=4)
    dm = _device_model(kind, n, obs, layers, S.Config(frequency=6), state, 4)
    assert (dm.n_terms, dm.n_monomials) == (ref.n_terms, ref.n_monomials)
    ps = random_params(n, 5, 4)
    assert list(dm.evaluate(np.array(ps))) == [ref.evaluate(p) for p in ps]


def _hea(n, depth, seed):
    """A brickwork of symbolic XX / ZZ couplings and X / Z fields."""
    rng = random.Random(seed)
    layers, p = [], 0
    for _ in range(depth):
        for q in range(n):
            layers.append([((1 << 2 * q), p, None)])  # X_q
            p += 1
        for q in range(n - 1):
            k = (1 << 2 * q) | (1 << 2 * (q + 1)) if rng.random() < 0.5 else (2 << 2 * q) | (2 << 2 * (q + 1))
            layers.append([(k, p, None)])
            p += 1
    return layers, p


def test_large_tape_matches_host_tape_evaluation():
    n, depth = 7, 3
    layers, n_params = _hea(n, depth, 3)
    obs = [((2 << 2 * q), 1.0 / n) for q in range(n)]  # sum_q Z_q / n
    cfg = S.Config(coefficient=1e-4)
    print("surrogate large tape", flush=True)
    dm = _device_model(S.PAULI, n, obs, layers, cfg, 0b1010, n_params)
    assert dm.n_ops > 1000
    kinds, params, a, b, vals, ov, roots = dm.export()
    roots = [-1 if r == np.uint64(2 ** 64 - 1) else int(r) for r in roots]
    ps = random_params(4, 6, n_params)
    want = S.evaluate_tape(list(map(int, kinds)), list(map(int, params)), list(map(int, a)), list(map(int, b)),
                           list(map(float, vals)), list(map(float, ov)), roots, ps)
    assert list(dm.evaluate(np.array(ps))) == want
    # many sets at once equal the same sets one by one, and in passes of 7 and 128 sets
    many = random_params(5, 300, n_params)
    batch = dm.evaluate(np.array(many))
    for i in (0, 17, 299):
        assert dm.evaluate(np.array(many[i:i + 1]))[0] == batch[i]
    for limit in (7, 128):
        dm.set_batch_limit(limit)
        assert list(dm.evaluate(np.array(many))) == list(batch)
    dm.set_batch_limit(0)


def test_device_model_from_exported_tape():
    from propaq_amd import _native

    obs, layers = random_case(71, S.PAULI, n=5)
    dm = _device_model(S.PAULI, 5, obs, layers, S.Config(), 3, 5)
    dm2 = _native.DeviceModel.from_tape(dm.n_params, *dm.export())
    ps = np.array(random_params(8, 5, 5))
    assert list(dm2.evaluate(ps)) == list(dm.evaluate(ps))
    assert (dm2.n_terms, dm2.n_ops, dm2.n_monomials) == (dm.n_terms, dm.n_ops, dm.n_monomials)


def test_model_create_rejects_bad_tapes():
    from propaq_amd import _native

    with pytest.raises(ValueError):  # operand after its op
        _native.DeviceModel.from_tape(1, [1, 0], [0, 0], [1, 0], [1, 0], [0.0, 1.0], [1.0], [0])
    with pytest.raises(ValueError):  # parameter index beyond n_params
        _native.DeviceModel.from_tape(1, [0, 3], [0, 1], [0, 0], [0, 0], [1.0, 0.0], [1.0], [1])
    with pytest.raises(ValueError):  # root beyond the tape
        _native.DeviceModel.from_tape(0, [0], [0], [0], [0], [1.0], [1.0], [5])


# ------------------------------------------------------------ Python API

def _pauli(x, z, n=4):
    import propaq_amd as pq

    return pq.PauliString(x, z, n)


def _numeric_ev(obs, circ, state=0):
    import propaq_amd as pq

    return pq.PauliPropagator().expectation_value(obs, circ, initial_state=state).expectation_value


def test_api_numerical_agreement():
    import propaq_amd as pq

    obs = pq.PauliTermSum({_pauli(0, 0b0001): 1.0, _pauli(0, 0b0010): 0.5})
    gens = [_pauli(0b0001, 0), _pauli(0b0011, 0), _pauli(0, 0b0010), _pauli(0b0100, 0b0001)]
    angles = [0.3, 0.7, 1.1, 0.25]
    circ = pq.PauliCircuit([pq.PauliRotation(g, a) for g, a in zip(gens, angles)])
    sc = pq.SurrogatePauliCircuit.from_pauli_circuit(circ, param_indices=[0, 1, 2, 0])
    model = pq.PauliSurrogatePropagator().build(obs, sc, initial_state=0)
    assert model.n_params == 3
    for state in (0, 1):
        model = pq.PauliSurrogatePropagator().build(obs, sc, initial_state=state)
        num = _numeric_ev(obs, pq.PauliCircuit([pq.PauliRotation(g, a) for g, a in
                                                zip(gens, [0.3, 0.7, 1.1, 0.3])]), state)
        assert model.evaluate([0.3, 0.7, 1.1]) == pytest.approx(num, rel=1e-9, abs=1e-12)
    # a numeric gate stays baked in (param_index None)
    sc2 = pq.SurrogatePauliCircuit.from_pauli_circuit(circ, param_indices=[0, None, 1, None])
    m2 = pq.PauliSurrogatePropagator().build(obs, sc2, initial_state=0)
    assert m2.evaluate([0.3, 1.1]) == pytest.approx(_numeric_ev(obs, circ), rel=1e-9, abs=1e-12)


def test_api_frequency_policy_and_truncators():
    import propaq_amd as pq

    obs = pq.PauliTermSum({_pauli(0, 0b0001): 1.0})
    gens = [_pauli(0b0001, 0), _pauli(0b0011, 0), _pauli(0, 0b0010)]
    angles = [0.3, 0.7, 1.1]
    circ = pq.PauliCircuit([pq.PauliRotation(g, a) for g, a in zip(gens, angles)])
    sc = pq.SurrogatePauliCircuit.from_pauli_circuit(circ, param_indices=[0, 1, 2])
    num = _numeric_ev(obs, circ)
    prev = math.inf
    for f in range(4):
        m = pq.PauliSurrogatePropagator(truncation=pq.FrequencyTruncationPolicy(max_frequency=f)).build(obs, sc)
        err = abs(m.evaluate(angles) - num)
        assert err <= prev + 1e-12
        prev = err
    assert prev < 1e-9
    m = pq.PauliSurrogatePropagator(truncation=[pq.CoefficientTruncator(1e9)]).build(obs, sc)
    assert m.evaluate(angles) == pytest.approx(0.0, abs=1e-12)
    m = pq.PauliSurrogatePropagator(truncation=[pq.FrequencyTruncator(None), pq.CoefficientTruncator(None),
                                                pq.WeightTruncator(None), pq.TermBudget(None)]).build(obs, sc)
    assert m.evaluate(angles) == pytest.approx(num, rel=1e-9)
    with pytest.raises(ValueError):
        pq.PauliSurrogatePropagator(truncation=pq.NativeTruncator.__new__(pq.NativeTruncator))
    prop = pq.PauliSurrogatePropagator(truncation=[pq.TermBudget(10), pq.WeightTruncator(2)])
    assert len(prop.truncators) == 2
    prop.set_truncation([pq.WeightTruncator(5)])
    assert len(prop.truncators) == 1


def test_api_evaluate_batch_and_save_load(tmp_path):
    import propaq_amd as pq

    obs = pq.PauliTermSum({_pauli(0, 0b0001): 1.0})
    rots = [pq.PauliRotation(_pauli(0b0001, 0), 0.0), pq.PauliRotation(_pauli(0b0010, 0b0001), 0.0),
            pq.PauliRotation(_pauli(0b0100, 0b0010), 0.0)]
    sc = pq.SurrogatePauliCircuit.from_pauli_circuit(pq.PauliCircuit(rots), param_indices=[0, 1, 2])
    model = pq.PauliSurrogatePropagator().build(obs, sc, initial_state=0)
    sets = [[0.1 * (k + 1), 0.2 * (k + 1), 0.3 * (k + 1)] for k in range(5)]
    batch = model.evaluate_batch(sets)
    assert batch == [model.evaluate(p) for p in sets]
    assert model.evaluate_batch([]) == []
    with pytest.raises(ValueError):
        model.evaluate_batch([[0.1, 0.2, 0.3], [0.1]])
    with pytest.raises(ValueError):
        model.evaluate([0.1])
    path = os.path.join(str(tmp_path), "m.surrogate.gz")
    model.save(path)
    loaded = pq.PauliSurrogateModel.load(path)
    assert loaded.evaluate_batch(sets) == batch
    assert (loaded.n_terms, loaded.n_params, loaded.n_monomials) == (model.n_terms, model.n_params,
                                                                     model.n_monomials)
    bad = os.path.join(str(tmp_path), "bad.gz")
    with open(bad, "wb") as fh:
        fh.write(b"not a surrogate model file")
    with pytest.raises(Exception):
        pq.PauliSurrogateModel.load(bad)


def test_api_majorana_matches_numeric():
    import propaq_amd as pq

    n_modes = 8
    obs = pq.MajoranaTermSum({pq.MajoranaMonomial(0b0011, n_modes): 1.0, pq.MajoranaMonomial(0b1100, n_modes): 0.5})
    gens = [pq.MajoranaMonomial(m, n_modes) for m in (0b0110, 0b1001, 0b0101_0000 | 0b0110, 0b0011_0000)]
    angles = [0.4, -0.9, 1.3, 0.2]
    circ = pq.MajoranaCircuit([pq.MajoranaRotation(g, a) for g, a in zip(gens, angles)])
    sc = pq.SurrogateMajoranaCircuit.from_majorana_circuit(circ, param_indices=[0, 1, 2, 3])
    for state in (0, 0b0101):
        m = pq.MajoranaSurrogatePropagator().build(obs, sc, initial_state=state)
        num = pq.MajoranaPropagator().expectation_value(obs, circ, initial_state=state).expectation_value
        assert m.evaluate(angles) == pytest.approx(num, rel=1e-9, abs=1e-12)


def test_api_build_logger(tmp_path):
    import propaq_amd as pq

    obs = pq.PauliTermSum({_pauli(0, 0b0001): 1.0, _pauli(0, 0b0100): -0.5})
    gens = [_pauli(0b0001, 0), _pauli(0b0011, 0), _pauli(0, 0b0010), _pauli(0b0110, 0b0001), _pauli(0b1000, 0)]
    circ = pq.PauliCircuit([[pq.PauliRotation(g, 0.1)] for g in gens[:2]] + [[pq.PauliRotation(g, 0.1)
                                                                             for g in gens[2:]]])
    sc = pq.SurrogatePauliCircuit.from_pauli_circuit(circ, param_indices=[0, 1, 2, 0, 1])
    path = os.path.join(str(tmp_path), "sur.jsonl")
    prop = pq.PauliSurrogatePropagator(truncation=pq.FrequencyTruncationPolicy(max_frequency=3),
                                       logger=pq.Logger(path, 1))
    model = prop.build(obs, sc, initial_state=0)
    lp = pq.LogParser(path)
    assert lp.terms == prop.last_n_terms
    assert lp.gate_indices == list(range(5))
    # execution order: the last layer (three gates) first
    assert [e.layer_idx for e in lp.gate_events] == [0, 0, 0, 1, 2]
    assert [e.frequency for e in lp.surrogate_merge_events] == [3] * 5
    assert lp.monomials_before[0] == 2
    assert lp.engine_phases_events[0].terms == prop.last_n_terms[-1]
    assert model.n_terms <= prop.last_n_terms[-1]


def test_api_empty_circuit_and_many_terms(tmp_path):
    """tests/surrogate/test_surrogate_pauli.py: an empty circuit evaluates to the
    observable's trace, and a 15-term model survives save / load."""
    import propaq_amd as pq

    obs = pq.PauliTermSum({_pauli(0, 0b0001): 1.0})
    sc = pq.SurrogatePauliCircuit.from_pauli_circuit(pq.PauliCircuit([]), param_indices=[])
    model = pq.PauliSurrogatePropagator().build(obs, sc, initial_state=0)
    assert model.n_params == 0
    assert model.evaluate([]) == pytest.approx(1.0)
    obs15 = pq.PauliTermSum({_pauli(0, z): 1.0 + 0.1 * z for z in range(1, 16)})
    m15 = pq.PauliSurrogatePropagator().build(obs15, sc, initial_state=0)
    assert m15.n_terms == 15
    path = os.path.join(str(tmp_path), "m15.gz")
    m15.save(path)
    loaded = pq.PauliSurrogateModel.load(path)
    assert loaded.n_terms == 15
    assert loaded.evaluate([]) == m15.evaluate([])
    # numeric-only circuit: no parameters, values baked in
    circ = pq.PauliCircuit([pq.PauliRotation(_pauli(0b0001, 0), 0.4), pq.PauliRotation(_pauli(0b0011, 0), -0.2)])
    scn = pq.SurrogatePauliCircuit.from_pauli_circuit(circ, param_indices=[None, None])
    mn = pq.PauliSurrogatePropagator().build(obs, scn, initial_state=0)
    assert mn.n_params == 0
    assert mn.evaluate([]) == pytest.approx(_numeric_ev(obs, circ), rel=1e-12, abs=1e-14)


def test_model_monomial_count_saturates():
    """model.rs tests (n_monomials_saturates_instead_of_wrapping...): 135
    doublings of one scalar in three terms count 2^128 - 1 monomials."""
    from propaq_amd import _native

    n = 136
    kinds = [0] + [1] * (n - 1)
    a = [0] + list(range(n - 1))
    dm = _native.DeviceModel.from_tape(1, kinds, [0] * n, a, a, [1.5] + [0.0] * (n - 1), [1.0, -0.5, 2.0],
                                       [n - 1, n - 1, n - 1])
    assert dm.n_monomials == 2 ** 128 - 1
    got = dm.evaluate(np.array([[0.3]]))[0]
    want = O.tape_eval(kinds, [0] * n, a, a, [1.5] + [0.0] * (n - 1), [1.0, -0.5, 2.0], [n - 1] * 3, 1, [[0.3]])[0]
    assert got == want == 1.5 * 2.0 ** 135 * 2.5
