"""The partitioned, threaded oracle (the multi-core CPU baseline) against the
single-partition oracle.

Restates the reference's own partition tests
(crates/core/tests/unit/engine/partitioned_termsum.rs), with its toy algebra
(anticommutation = parity of the shared bits, product = XOR with phase i,
weight = bit count) and its xorshift seeds, then checks bit-exact agreement
on Pauli and Majorana circuits at one, two and four key words: per-gate live
counts, every term's coefficient, the integer counters."""

import math

import numpy as np
import pytest

from oracle import oracle as orc
from propaq_amd import runner, workloads as wl
from tests.parity import random_case

TOY = orc.TOY


class Rng:
    """partitioned_termsum.rs:58-74 (xorshift64)."""

    def __init__(self, s):
        self.s = s & 0xFFFFFFFFFFFFFFFF

    def next_u64(self):
        x = self.s
        x ^= (x << 13) & 0xFFFFFFFFFFFFFFFF
        x ^= x >> 7
        x ^= (x << 17) & 0xFFFFFFFFFFFFFFFF
        self.s = x
        return x

    def below(self, n):
        return self.next_u64() % n

    def unit(self):
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))


def mono(*bits):
    k = 0
    for b in bits:
        k |= 1 << b
    return k


def _engines(partitions, min_coeff=None):
    single = orc.OracleEngine(TOY, 8, 1)
    part = orc.PartitionedOracle(TOY, 8, 1, partitions, min(partitions, 4))
    for e in (single, part):
        e.set_cutoff(None, min_coeff, None)
        e.set_defer(False)
    return single, part


def _run_both(seed, partitions, n_gates, min_coeff=None):
    """run_both_inner (partitioned_termsum.rs:100-140)."""
    rng = Rng(seed)
    seeds = [(mono(rng.below(6)), 1.0 + rng.unit()) for _ in range(4)]
    gates = []
    for _ in range(n_gates):
        a = rng.below(6)
        b = (a + 1 + rng.below(5)) % 6
        gates.append((mono(a, b), 0.1 + rng.unit()))
    single, part = _engines(partitions, min_coeff)
    for e in (single, part):
        for k, c in seeds:
            e.add(k, c)
        for g, a in gates:
            e.apply_rotation(g, a)
    nz = lambda t: {k: v for k, v in t.items() if v != 0.0}
    return nz(single.terms()), nz(part.terms())


def test_one_partition_matches_the_single_partition_engine():
    want, got = _run_both(0x9E3779B97F4A7C15, 1, 20)
    assert got == want


@pytest.mark.parametrize("s", [1, 2, 3, 5, 8])
def test_the_pair_rescue_is_independent_of_partition_count(s):
    want, got = _run_both(0x123456789ABCDEF1, s, 40, min_coeff=0.1)
    assert got == want  # the reference allows 1e-9; per key the arithmetic is the same, so the bits agree


@pytest.mark.parametrize("s", [1, 2, 3, 4, 8, 16])
def test_partition_count_does_not_change_the_result(s):
    want, got = _run_both(0x2545F4914F6CDD1D, s, 24)
    assert got == want


def test_reclaim_drops_decayed_terms_and_leaves_the_store_usable():
    op = orc.PartitionedOracle(TOY, 8, 1, 4, 4)
    op.set_cutoff(None, 0.1, None)
    op.set_defer(False)
    for k in range(24):
        op.add(mono(k % 6, k // 6 + 1), 1.0)
    before = len(op)
    assert before > 4
    op.scale_by_table([1e-6] * 9)
    assert op.reclaim() == before
    assert len(op) == 0
    op.set_cutoff(None, None, None)
    op.add(mono(0), 1.0)
    op.apply_rotation(mono(0), 0.3)
    assert len(op) == 2, "the rebuilt store must still branch"


def test_reclaim_keeps_what_the_cutoff_still_admits():
    op = orc.PartitionedOracle(TOY, 8, 1, 4, 4)
    op.set_cutoff(None, 0.1, None)
    op.add(mono(0), 1.0)
    op.add(mono(1), 1e-9)
    op.add(mono(2), 0.5)
    assert op.reclaim() == 1
    assert len(op) == 2


def test_reclaim_without_a_cutoff_is_a_no_op():
    op = orc.PartitionedOracle(TOY, 8, 1, 4, 4)
    op.add(mono(0), 1e-30)
    assert op.reclaim() == 0
    assert len(op) == 1


def test_a_term_lives_only_in_the_partition_that_owns_its_key():
    op = orc.PartitionedOracle(TOY, 8, 1, 4, 4)
    op.set_defer(False)
    rng = Rng(0x853C49E6748FEA9B)
    for _ in range(64):
        op.add(mono(rng.below(6)), 1.0)
    for _ in range(12):
        a = rng.below(6)
        op.apply_rotation(mono(a, (a + 1) % 6), 0.3)
    keys, _ = op.export()
    lens = op.partition_lens()
    assert sum(lens) == len(keys)
    r = 0
    for q, n in enumerate(lens):
        for k in keys[r:r + n]:
            assert op.owner(orc.from_words(k)) == q
        r += n
    # ownership is (hash >> 32) mod S with splitmix64 of a one-word key (termsum.rs:872-877)
    assert len(set(lens)) > 1 or max(lens) > 0


def test_expectation_agrees_across_partition_counts():
    base = None
    for s in (1, 2, 5, 8):
        op = orc.PartitionedOracle(TOY, 8, 1, s, min(s, 4))
        op.set_defer(False)
        rng = Rng(0xD1B54A32D192ED03)
        for _ in range(16):
            op.add(mono(rng.below(6)), 1.0 + rng.unit())
        for _ in range(10):
            a = rng.below(6)
            op.apply_rotation(mono(a, (a + 1) % 6), 0.3)
        v = op.expectation(0b101)
        if base is None:
            base = v
        assert abs(v - base) < 1e-9


def _check_case(kind, n_units, obs, layers, s, **kw):
    one = orc.run_oracle(kind, n_units, obs, layers, want_terms=True, **kw)
    part = orc.run_oracle(kind, n_units, obs, layers, want_terms=True, partitions=s, threads=min(s, 4), **kw)
    assert part.n_terms == one.n_terms
    assert part.terms == one.terms
    for k in ("visited", "emitted", "declined", "hits"):
        assert part.counters[k] == one.counters[k], k
    assert math.isclose(part.counters["discarded_coeff_l1"], one.counters["discarded_coeff_l1"], rel_tol=1e-12,
                        abs_tol=1e-300)
    assert abs(part.expectation_value - one.expectation_value) <= 1e-12
    assert part.terms_below_cutoff == one.terms_below_cutoff


@pytest.mark.parametrize("s", [2, 3, 4, 8])
@pytest.mark.parametrize("kind,n_units", [(0, 20), (0, 40), (0, 100), (1, 24), (1, 60)])
def test_random_truncated_noisy_circuits_bit_exact(kind, n_units, s):
    obs, layers, diag = random_case(700 + n_units, kind, n_units, n_layers=8, gates_per_layer=12)
    tbl = runner.uniform_table(0.03, n_units)
    _check_case(kind, n_units, obs, layers, s, max_weight=5, min_coeff=1e-3, noise_table=lambda li, nl: tbl,
                diag_state=diag)


@pytest.mark.parametrize("s", [3, 8])
def test_kicked_ising_2d_prefix_bit_exact(s):
    # W = 4, deferral off (RZZ(-pi/2) changes weight), noise and reclaim: 6.8e4 rows
    w = wl.config(2)
    tbl = w.noise_table()
    _check_case(w.kind, w.n_units, w.observable, w.layers[:6], s, max_weight=w.max_weight, min_coeff=w.min_coeff,
                noise_table=lambda li, nl: tbl, diag_state=w.diag_state)


@pytest.mark.parametrize("s", [4])
def test_clifford_frame_shared_by_partitions(s):
    # no weight cutoff: Clifford gates fold into the one shared frame
    obs, layers, diag = random_case(741, 0, 16, n_layers=8, gates_per_layer=10, clifford_frac=0.5)
    _check_case(0, 16, obs, layers, s, min_coeff=1e-3, diag_state=diag)
