"""Pins of the surrogate oracle (oracle/surrogate.py) and checks of the
native builder's host logic against it.  CPU only: the builder runs on the
host; compiling and evaluating a model needs the device (test_gpu_surrogate).

What pins the oracle, so that a plausible mistake fails one of these:
* its textbook algebra equals the C oracle's (itself pinned against the
  reference's fixtures) on random keys;
* with no truncation, every term's coefficient evaluated at angles theta
  equals, bit for bit, the C oracle's numeric propagation of the same circuit
  bound to theta, key for key (frames and signs included);
* frequency pruning keeps exactly the monomials with at most f trig factors,
  by brute-force expansion of the unpruned coefficients;
* coefficient pruning drops only monomials below the cutoff;
* simplify leaves every value unchanged and its monomials distinct;
* the reference's own known answers (crates/surrogate tests,
  tests/surrogate/test_surrogate_pauli.py): echo, empty circuit, exactness at
  full frequency, monotone error in the frequency cap.
"""

from __future__ import annotations

import math
import os
import random

import numpy as np
import pytest

from oracle import oracle as O
from oracle import surrogate as S
from surrogate_cases import CONFIGS, numeric_layers, random_case, random_params, wide_case


def test_algebra_matches_c_oracle():
    rng = random.Random(11)
    for kind in (S.PAULI, S.MAJORANA):
        n = 5
        for _ in range(200):
            a, b = rng.getrandbits(2 * n), rng.getrandbits(2 * n)
            key, e = O.product(kind, a, b, n)
            assert S.product(kind, a, b, n) == (e % 4, key)
            assert S.anticommutes(kind, a, b, n) == O.anticommutes(kind, a, b, n)
            st = rng.getrandbits(n)
            assert S.basis_trace(kind, a, n, st) == O.trace(kind, a, n, st)
            assert S.weight(kind, a, n) == O.weight(kind, a, n)


@pytest.mark.parametrize("kind", [S.PAULI, S.MAJORANA])
@pytest.mark.parametrize("seed", range(6))
def test_coefficients_equal_numeric_propagation(kind, seed):
    obs, layers = random_case(100 + seed, kind)
    terms, frame = S.propagate(kind, 4, obs, layers, S.Config())
    for ps in random_params(seed, 2, 5):
        model = S.Model([(k, 1.0, terms[k]) for k in sorted(terms)], 5)
        vals = model.coefficient_values(ps)
        logical = {}
        for (k, _, _), v in zip(model.terms, vals):
            img, sign = frame.readout(k)
            logical[img] = sign * v
        ref = O.run_oracle(kind, 4, obs, numeric_layers(layers, ps), want_terms=True).terms
        assert logical.keys() == ref.keys()
        for k, v in ref.items():
            assert logical[k] == v, (hex(k), logical[k], v)


@pytest.mark.parametrize("kind", [S.PAULI, S.MAJORANA])
@pytest.mark.parametrize("f", [0, 1, 2, 4])
def test_frequency_prune_is_brute_force(kind, f):
    obs, layers = random_case(7 + f, kind, clifford=0.0)
    full, _ = S.propagate(kind, 4, obs, layers, S.Config())
    cut, _ = S.propagate(kind, 4, obs, layers, S.Config(frequency=f))
    ps = random_params(f, 1, 5)[0]
    lut = {p: (math.cos(t), math.sin(t)) for p, t in enumerate(ps)}

    def value(monos):
        tot = 0.0
        for facs, s in monos:
            v = s
            for p, k in facs:
                v *= lut[p][0] if k == S.COS else lut[p][1]
            tot += v
        return tot

    for k, root in full.items():
        keep = [m for m in S.monomials(root) if len(m[0]) <= f]
        if not keep:
            assert k not in cut
            continue
        got = S.monomials(cut[k])
        assert sorted(got) == sorted(keep)
        assert cut[k].count == len(keep)
        assert S.Model([(k, 1.0, cut[k])], 5).coefficient_values(ps)[0] == pytest.approx(value(keep), rel=1e-12,
                                                                                          abs=1e-15)
    assert set(cut) <= set(full)


@pytest.mark.parametrize("cutoff", [0.02, 0.1, 0.3])
def test_coefficient_prune_drops_only_small_monomials(cutoff):
    obs, layers = random_case(31, S.PAULI, numeric=0.3, clifford=0.1)
    full, _ = S.propagate(S.PAULI, 4, obs, layers, S.Config())
    cut, _ = S.propagate(S.PAULI, 4, obs, layers, S.Config(coefficient=cutoff))
    n_dropped = 0
    for k, root in full.items():
        have = S.monomials(root)
        kept = S.monomials(cut.get(k))
        rest = list(have)
        for m in kept:
            rest.remove(m)  # kept is a sub-multiset of the unpruned monomials
        for _, s in rest:
            assert abs(s) < cutoff
        n_dropped += len(rest)
    assert n_dropped > 0


@pytest.mark.parametrize("kind", [S.PAULI, S.MAJORANA])
def test_simplify_keeps_values(kind):
    obs, layers = random_case(41, kind, n_params=2)
    plain, _ = S.propagate(kind, 4, obs, layers, S.Config())
    simp, _ = S.propagate(kind, 4, obs, layers, S.Config(simplify=True))
    assert plain.keys() == simp.keys()
    ps = random_params(3, 1, 2)[0]
    for k in plain:
        a = S.Model([(k, 1.0, plain[k])], 2).coefficient_values(ps)[0]
        b = S.Model([(k, 1.0, simp[k])], 2).coefficient_values(ps)[0]
        assert b == pytest.approx(a, rel=1e-12, abs=1e-14)
        runs = [f for f, _ in S.expand(simp[k])]
        assert len(runs) == len(set(runs))
        assert simp[k].count <= plain[k].count


def _pkey(x, z, n=4):
    k = 0
    for q in range(n):
        k |= ((x >> q) & 1) << (2 * q)
        k |= ((z >> q) & 1) << (2 * q + 1)
    return k


def test_reference_known_answers():
    z0 = _pkey(0, 0b0001)
    x0 = _pkey(0b0001, 0)
    # empty circuit: the observable's own trace (test_surrogate_pauli.py:128)
    m = S.build(S.PAULI, 4, [(z0, 1.0)], [], S.Config(), 0, 0)
    assert m.evaluate([]) == 1.0
    # Loschmidt echo U(theta) then U(-theta): <0|Z|0> = 1 (:338)
    m = S.build(S.PAULI, 4, [(z0, 1.0)], [[(x0, 0, None)], [(x0, 1, None)]], S.Config(), 0, 2)
    assert m.evaluate([0.9, -0.9]) == pytest.approx(1.0, abs=1e-12)
    # one rotation: cos(theta) on |0>, -cos(theta) on |1> (:114)
    m0 = S.build(S.PAULI, 4, [(z0, 1.0)], [[(x0, 0, None)]], S.Config(), 0, 1)
    m1 = S.build(S.PAULI, 4, [(z0, 1.0)], [[(x0, 0, None)]], S.Config(), 1, 1)
    assert m0.evaluate([0.4]) == math.cos(0.4)
    assert m1.evaluate([0.4]) == -math.cos(0.4)
    # frequency caps: exact at the number of rotations, error monotone (:166-190)
    gens = [_pkey(0b0001, 0), _pkey(0b0011, 0), _pkey(0, 0b0010)]
    angles = [0.3, 0.7, 1.1]
    layers = [[(g, i, None)] for i, g in enumerate(gens)]
    num = O.run_oracle(S.PAULI, 4, [(z0, 1.0)], [[(g, a)] for g, a in zip(gens, angles)]).expectation_value
    prev = math.inf
    for f in range(4):
        v = S.build(S.PAULI, 4, [(z0, 1.0)], layers, S.Config(frequency=f), 0, 3).evaluate(angles)
        assert abs(v - num) <= prev + 1e-12
        prev = abs(v - num)
    assert prev < 1e-12
    # a coefficient cutoff above every scalar prunes everything (:276)
    m = S.build(S.PAULI, 4, [(z0, 1.0)], layers, S.Config(coefficient=1e9), 0, 3)
    assert m.evaluate(angles) == 0.0


def test_tape_evaluation_matches_dag():
    for kind in (S.PAULI, S.MAJORANA):
        obs, layers = random_case(77, kind)
        m = S.build(kind, 4, obs, layers, S.Config(), 5, 5)
        ps = random_params(9, 4, 5)
        assert S.evaluate_tape(*S.compile_tape(m), ps) == [m.evaluate(p) for p in ps]


def test_pairwise_sum_tree():
    assert S.pairwise_sum([]) == 0.0
    v = [1e16, 1.0, -1e16, 1.0]
    assert S.pairwise_sum(v) == (1e16 + 1.0) + (-1e16 + 1.0)
    w = [0.1 * i for i in range(37)]
    # padding to any power of two gives the same tree
    assert S.pairwise_sum(w) == S.pairwise_sum(w + [0.0] * 27)


# ----------------------------------------------- native builder (host only)

def _builder_run(kind, n, obs, layers, cfg, compact_min=None):
    from propaq_amd import _native
    from propaq_amd.runner import keys_to_array

    b = _native.SurrogateBuilder(kind, n)
    if compact_min is not None:
        b.set_compaction(compact_min)
    b.set_truncation(frequency=cfg.frequency, coefficient=cfg.coefficient, weight=cfg.weight,
                     min_terms=cfg.min_terms, simplify=cfg.simplify)
    b.seed(keys_to_array([k for k, _ in obs], b.W), np.array([c for _, c in obs]))
    gates = [g for layer in reversed(layers) for g in reversed(layer)]
    gens = keys_to_array([g for g, _, _ in gates], b.W)
    pidx = np.array([-1 if p is None else p for _, p, _ in gates], dtype=np.int64)
    ang = np.array([0.0 if a is None else a for _, _, a in gates])
    counts, monos = b.apply(gens, pidx, ang, monomials=True)
    n_live = len(b)
    _builder_run.nodes = b.counters()["nodes"]
    b.close()
    return [int(c) for c in counts], monos, n_live


@pytest.mark.parametrize("seed", range(28))
def test_builder_counts_match_oracle(seed):
    """Odd seeds sweep unreachable nodes out of the arena after every pass
    that dropped something (compaction from 16 nodes on)."""
    kind = seed % 2
    cfg = CONFIGS[seed % len(CONFIGS)]
    obs, layers = random_case(500 + seed, kind, n=4 + seed % 3, n_gates=14)
    counts, monos = [], []
    S.propagate(kind, 4 + seed % 3, obs, layers, cfg, counts, monos)
    got_c, got_m, n_live = _builder_run(kind, 4 + seed % 3, obs, layers, cfg,
                                        compact_min=16 if seed % 2 else None)
    assert got_c == counts
    assert got_m == monos
    assert n_live == (counts[-1] if counts else len(obs))


@pytest.mark.parametrize("kind", [S.PAULI, S.MAJORANA])
@pytest.mark.parametrize("n", [40, 100, 200])
def test_builder_counts_match_oracle_wide_keys(kind, n):
    """Key widths W = 2, 4, 8 (n_units 40, 100, 200): sparse generators that
    touch units across every key word."""
    obs, layers = wide_case(kind, n, n + kind)
    cfg = S.Config(frequency=6)
    counts, monos = [], []
    S.propagate(kind, n, obs, layers, cfg, counts, monos)
    got_c, got_m, _ = _builder_run(kind, n, obs, layers, cfg)
    assert got_c == counts
    assert got_m == monos


def test_builder_rejects_bad_arguments():
    from propaq_amd import _native

    with pytest.raises(ValueError):
        _native.SurrogateBuilder(2, 4)
    with pytest.raises(ValueError):
        _native.SurrogateBuilder(0, 300)


def test_c_tape_evaluation_matches_python():
    for kind in (S.PAULI, S.MAJORANA):
        obs, layers = random_case(88, kind, n=5)
        m = S.build(kind, 5, obs, layers, S.Config(coefficient=0.01), 3, 5)
        tape = S.compile_tape(m)
        ps = random_params(12, 5, 5)
        assert list(O.tape_eval(*tape, 5, ps)) == S.evaluate_tape(*tape, ps)
    # empty model and empty coefficients
    assert list(O.tape_eval([], [], [], [], [], [], [], 0, [[0.0]])) == [0.0]
    assert list(O.tape_eval([0], [0], [0], [0], [2.0], [0.5, 3.0], [0, -1], 0, [[0.0]])) == [1.0]


def test_cos_and_sin_are_not_fused():
    """cos and sin of each angle are separate C library calls in the engine and
    the oracle alike: glibc's sincos differs from them in the last bit for
    about one angle in a thousand, so a fused call would break bit parity."""
    import shutil
    import subprocess

    if shutil.which("nm") is None:
        pytest.skip("nm not available")
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for lib in (os.path.join(here, "propaq_amd", "_lib", "libpropaq_amd.so"),
                os.path.join(here, "oracle", "_lib", "libpq_oracle.so")):
        if not os.path.exists(lib):
            continue
        syms = subprocess.run(["nm", "-D", lib], capture_output=True, text=True).stdout
        assert " U sincos" not in syms, lib


def test_builder_log_records(tmp_path):
    """The logged build's per-gate records (gate, surrogate_merge) against the
    oracle's per-gate counts, through the log writer and LogParser."""
    from propaq_amd import _native
    from propaq_amd.events import LogParser, Logger, phases_record, write_surrogate_log
    from propaq_amd.runner import keys_to_array

    obs, layers = random_case(321, S.PAULI, n=5, n_gates=12)
    cfg = S.Config(frequency=3, weight=3)
    counts, monos = [], []
    S.propagate(S.PAULI, 5, obs, layers, cfg, counts, monos)
    b = _native.SurrogateBuilder(S.PAULI, 5)
    b.set_truncation(frequency=3, weight=3)
    b.seed(keys_to_array([k for k, _ in obs], b.W), np.array([c for _, c in obs]))
    gates = [g for layer in reversed(layers) for g in reversed(layer)]
    stats = b.apply_logged(keys_to_array([g for g, _, _ in gates], b.W),
                           np.array([-1 if p is None else p for _, p, _ in gates], dtype=np.int64),
                           np.array([0.0 if a is None else a for _, _, a in gates]))
    assert [st["terms"] for st in stats] == counts
    assert [st["monomials"] for st in stats] == monos
    assert all(st["terms_rotated"] >= st["terms"] for st in stats)
    recs = [dict(st, gate_idx=i, layer_idx=i, qiskit_gate_idx=None) for i, st in enumerate(stats)]
    path = os.path.join(str(tmp_path), "s.jsonl")
    write_surrogate_log(Logger(path, 2), recs, {"frequency": 3, "weight": 3, "coefficient": None}, len(obs),
                        phases_record(counts[-1], b.W, {}, b.counters()))
    lp = LogParser(path)
    assert lp.gate_indices == list(range(0, 12, 2))
    assert lp.terms == counts[::2]
    assert lp.monomials_after == monos[::2]
    prev = [len(obs)] + monos[:-1]
    assert lp.monomials_before == prev[::2]
    assert [e.frequency for e in lp.surrogate_merge_events] == [3] * 6
    assert lp.engine_phases_events[0].terms == counts[-1]
    b.close()


def test_variational_binding():
    """VariationalSurrogateModel's binding of named parameters to raw slots
    (the behaviour of propaq/models/variational.py:39-55) on a stand-in model
    that records what it is asked to evaluate."""
    import propaq_amd as pq

    class Rec:
        def evaluate(self, p):
            return sum(p)

        def evaluate_batch(self, ps):
            return [sum(p) for p in ps]

    theta, phi = "theta", "phi"
    srcs = [pq.ParamSource(theta, 1.0), pq.ParamSource(phi, -2.0), pq.ParamSource(None, 0.5), pq.ParamSource(theta, 0.5)]
    v = pq.VariationalSurrogateModel(Rec(), srcs, (theta, phi))
    assert v.n_params == 4
    assert v.raw_vector({theta: 0.3, phi: 1.0}) == [0.3, -2.0, 0.5, 0.15]
    assert v.raw_vector([0.3, 1.0]) == v.raw_vector({theta: 0.3, phi: 1.0})
    assert v.evaluate_batch([[0.3, 1.0], {theta: 1.0, phi: 0.0}]) == [0.3 - 2.0 + 0.5 + 0.15, 2.0]
    with pytest.raises(ValueError):
        v.evaluate([0.1])


def test_compile_fails_loudly_without_a_device():
    """The model lives on the device: without one, compile and model creation
    raise instead of falling back to the host."""
    from propaq_amd import _native

    if _native.device_count() > 0:
        pytest.skip("a device is present")
    b = _native.SurrogateBuilder(S.PAULI, 4)
    b.seed(np.array([[2]], dtype=np.uint64), np.array([1.0]))
    with pytest.raises(RuntimeError):
        b.compile(0, 0)
    with pytest.raises(RuntimeError):
        _native.DeviceModel.from_tape(0, [0], [0], [0], [0], [1.0], [1.0], [0])
    b.close()


@pytest.mark.parametrize("simplify", [False, True])
def test_builder_deep_chain_on_a_small_stack(simplify):
    """A coefficient touched by every gate is a DAG as deep as the circuit
    (crates/surrogate tests: *_deep_chain_does_not_overflow_the_stack).  The
    builder's prune and simplify walks keep their own stacks: 4000 gates run
    in a thread with a 256 KiB stack, which one call frame per node would
    overflow."""
    import threading

    z0 = _pkey(0, 0b01, 2)
    gen = _pkey(0b01, 0b10, 2)  # X0 Z1: its branch of Z0 has weight 2 and is dropped
    n_g = 4000
    layers = [[(gen, i % 7, None)] for i in range(n_g)]
    cfg = S.Config(coefficient=1e-300, weight=1, simplify=simplify)
    out = {}

    def run():
        try:
            out["r"] = _builder_run(S.PAULI, 2, [(z0, 1.0)], layers, cfg)
        except BaseException as e:  # noqa: BLE001
            out["e"] = e

    old = threading.stack_size(256 * 1024)
    try:
        t = threading.Thread(target=run)
        t.start()
        t.join()
    finally:
        threading.stack_size(old)
    assert "e" not in out, out.get("e")
    counts, monos, n_live = out["r"]
    assert counts == [1] * n_g and n_live == 1
    assert monos == [1] * n_g


def _val(root, params):
    return S.Model([(0, 1.0, root)], len(params)).coefficient_values(params)[0]


def test_reference_node_level_known_answers():
    """crates/surrogate/tests/unit/symcoeff.rs, restated on the oracle's nodes."""
    t0, t1 = 0.37, 0.74
    # prune_hand_built_cross_check_frequency_and_coefficient (:692)
    m1 = S.n_trig(S.COS, 0, S.n_scale(2.0, S.n_scalar(3.0)))
    m2 = S.n_trig(S.SIN, 1, S.n_scalar(0.5))
    total = S.n_add(m1, m2)
    want = 6.0 * math.cos(t0) + 0.5 * math.sin(t1)
    assert _val(total, [t0, t1]) == pytest.approx(want, abs=1e-12)
    assert S.prune(total, 0, None) is None
    assert _val(S.prune(total, 1, None), [t0, t1]) == pytest.approx(want, abs=1e-12)
    assert _val(S.prune(total, None, 1.0), [t0, t1]) == pytest.approx(6.0 * math.cos(t0), abs=1e-12)
    assert S.prune(total, None, 10.0) is None
    # prune_with_no_cutoffs_is_a_true_no_op (:683)
    assert S.prune(total, None, None) is total
    # count_saturates_instead_of_wrapping_past_u128_max (:24)
    big = S.n_scalar(1.0)
    for _ in range(130):
        big = S.n_add(big, big)
    assert big.count == S.U128_MAX
    # simplify_drops_exact_cancellation_to_empty (:183)
    a = S.n_trig(S.COS, 0, S.n_scalar(3.0))
    b = S.n_trig(S.COS, 0, S.n_scalar(-3.0))
    assert S.simplify(S.n_add(a, b)) is None
    # same_parameter_at_two_gates_collapses_to_a_power (:115): cos^2 of one parameter, one monomial
    c2 = S.n_trig(S.COS, 0, S.n_trig(S.COS, 0, S.n_scalar(1.0)))
    assert S.expand(c2) == [(((0, (2, 0)),), 1.0)]
    # two_derivation_paths_through_the_same_parameter_sum_correctly (:126) + simplify collapses them (:148)
    p1 = S.n_trig(S.SIN, 1, S.n_trig(S.COS, 0, S.n_scalar(2.0)))
    p2 = S.n_trig(S.COS, 0, S.n_trig(S.SIN, 1, S.n_scalar(0.5)))
    s = S.simplify(S.n_add(p1, p2))
    assert s.count == 1
    assert _val(s, [t0, t1]) == pytest.approx(2.5 * math.cos(t0) * math.sin(t1), rel=1e-12)


def test_compaction_drops_unreachable_nodes():
    """With a small threshold, unreachable coefficient nodes are swept out of
    the arena; the per-gate counts are those of the unswept build."""
    obs, layers = random_case(77, S.PAULI, n=6, n_gates=20)
    cfg = S.Config(frequency=3, simplify=True)
    swept = _builder_run(S.PAULI, 6, obs, layers, cfg, compact_min=16)
    n_swept = _builder_run.nodes
    kept = _builder_run(S.PAULI, 6, obs, layers, cfg)
    n_kept = _builder_run.nodes
    assert swept == kept
    assert n_swept < n_kept


def test_surrogate_circuit_construction():
    """tests/surrogate/test_surrogate_majorana.py TestCircuitConstruction /
    TestNumericAngleRotations, and the Pauli twins: host-side only."""
    import propaq_amd as pq

    n_modes = 8
    mm = lambda m: pq.MajoranaMonomial(m, n_modes)  # noqa: E731
    gens = [mm(0b0110), mm(0b1001)]
    sc = pq.SurrogateMajoranaCircuit.from_generators_and_param_indices(gens, [0, 1], n_modes)
    assert (sc.n_params, len(sc.rotations), sc.n_modes) == (2, 2, n_modes)
    assert pq.SurrogateMajoranaCircuit.from_generators_and_param_indices(gens, [0, 0], n_modes).n_params == 1
    circ = pq.MajoranaCircuit([pq.MajoranaRotation(mm(0b0110), 0.3)], n_modes=n_modes)
    with pytest.raises(ValueError):
        pq.SurrogateMajoranaCircuit.from_majorana_circuit(circ, param_indices=[0, 1])
    g = mm(0b0110)
    layers = [[pq.SurrogateMajoranaRotation(g, angle=0.1)], [pq.SurrogateMajoranaRotation(g, param_index=0)],
              [pq.SurrogateMajoranaRotation(g, angle=0.2)], [pq.SurrogateMajoranaRotation(g, param_index=2)]]
    assert pq.SurrogateMajoranaCircuit(layers, n_modes).n_params == 3
    assert pq.SurrogateMajoranaCircuit([[pq.SurrogateMajoranaRotation(g, angle=0.1)]], n_modes).n_params == 0
    circ2 = pq.MajoranaCircuit([pq.MajoranaRotation(x, a) for x, a in zip(gens, [0.3, 0.6])], n_modes=n_modes)
    sc2 = pq.SurrogateMajoranaCircuit.from_majorana_circuit(circ2, param_indices=[None, 0])
    assert (sc2.rotations[0].param_index, sc2.rotations[0].angle) == (None, 0.3)
    assert (sc2.rotations[1].param_index, sc2.rotations[1].angle) == (0, None)
    for kw in ({}, {"param_index": 0, "angle": 0.5}, {"param_index": 0, "angle": 0.0}):
        with pytest.raises(ValueError):
            pq.SurrogateMajoranaRotation(g, **kw)
    ps = pq.PauliString(0b01, 0, 2)
    psc = pq.SurrogatePauliCircuit.from_pauli_circuit(pq.PauliCircuit([pq.PauliRotation(ps, 0.2)] * 3),
                                                      param_indices=[1, None, 4])
    assert psc.n_params == 5 and psc.rotations[1].angle == 0.2
