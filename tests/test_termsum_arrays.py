"""Array-backed term sums: the vectorized weights, file layout, truncation,
damping and merge agree with the per-term definitions (PauliString.weight,
jw_weight, to_bytes, TruncationPolicy.should_truncate, add())."""

import gzip
import random
import struct

import numpy as np
import pytest

from propaq_amd.datatypes import (MajoranaMonomial, MajoranaTermStreamer, MajoranaTermSum, PauliString,
                                  PauliTermStreamer, PauliTermSum, jw_weight)
from propaq_amd.truncation import TruncationPolicy


def _pauli_sum(n, count, seed):
    rng = random.Random(seed)
    ts = PauliTermSum()
    for _ in range(count):
        ts.add(PauliString(rng.getrandbits(n), rng.getrandbits(n), n), rng.uniform(-1, 1))
    return ts


def _majorana_sum(n, count, seed):
    rng = random.Random(seed)
    ts = MajoranaTermSum()
    for _ in range(count):
        ts.add(MajoranaMonomial(rng.getrandbits(n), n, False), rng.uniform(-1, 1))
    return ts


@pytest.mark.parametrize("n", [1, 7, 31, 32, 33, 64, 65, 200])
def test_pauli_weights_match_terms(n):
    ts = _pauli_sum(n, 200, n)
    assert list(ts.weights()) == [t.weight for t, _ in ts.items()]


@pytest.mark.parametrize("n", [2, 30, 62, 64, 66, 128, 130, 256])
def test_majorana_weights_match_jordan_wigner(n):
    ts = _majorana_sum(n, 200, n)
    assert list(ts.weights()) == [jw_weight(t.modes, n // 2) for t, _ in ts.items()]


def test_majorana_weight_closed_forms():
    # a single gamma at site s carries a Z string on every site below it: weight s + 1
    n_sites = 40
    for s in range(n_sites):
        for m in (2 * s, 2 * s + 1):
            t = MajoranaTermSum()
            t.add(MajoranaMonomial(1 << m, 2 * n_sites), 1.0)
            assert list(t.weights()) == [s + 1]
    # a number operator gamma_2s gamma_2s+1 is a single Z: weight 1
    t = MajoranaTermSum()
    t.add(MajoranaMonomial(3 << 70, 2 * n_sites), 1.0)
    assert list(t.weights()) == [1]


@pytest.mark.parametrize("make,n", [(_pauli_sum, 5), (_pauli_sum, 70), (_majorana_sum, 6), (_majorana_sum, 130)])
def test_file_layout_matches_to_bytes(tmp_path, make, n):
    ts = make(n, 100, 3)
    f = tmp_path / "t.gz"
    ts.save(str(f))
    with gzip.open(f, "rb") as h:
        count, stride, size = struct.unpack("<QQQ", h.read(24))
        assert (count, size) == (len(ts), n)
        for t, c in ts.items():
            assert h.read(stride) == t.to_bytes()
            assert struct.unpack("<d", h.read(8))[0] == c
    back = type(ts).from_file(str(f))
    assert back._terms == ts._terms and list(back._terms) == list(ts._terms) and back.n_units == n
    streamer = PauliTermStreamer if isinstance(ts, PauliTermSum) else MajoranaTermStreamer
    assert [(t.key, c) for t, c in streamer.from_file(str(f))] == [(t.key, c) for t, c in ts.items()]


def test_from_file_folds_repeated_keys(tmp_path):
    p = PauliString(5, 2, 4)
    q = PauliString(1, 1, 4)
    f = tmp_path / "r.gz"
    with gzip.open(f, "wb") as h:
        h.write(struct.pack("<QQQ", 3, 2, 4))
        for t, c in ((p, 0.1), (q, 0.2), (p, 0.3)):
            h.write(t.to_bytes() + struct.pack("<d", c))
    ts = PauliTermSum.from_file(str(f))
    assert ts._terms == {p.key: 0.1 + 0.3, q.key: 0.2}


def test_merge_matches_sequential_add():
    a = _pauli_sum(9, 80, 1)
    b = _pauli_sum(9, 80, 2)
    want = a.copy()
    for t, c in b.items():
        want.add(t, c)
    a.merge(b)
    assert list(a._terms.items()) == list(want._terms.items())


def test_merge_float32_rounds_each_add():
    a = PauliTermSum(dtype="float32")
    b = PauliTermSum(dtype="float32")
    p = PauliString(3, 0, 3)
    a.add(p, 0.1)
    b.add(p, 0.2)
    want = float(np.float32(float(np.float32(0.1)) + float(np.float32(0.2))))
    a.merge(b)
    assert a[p] == want
    with pytest.raises(ValueError):
        a.merge(PauliTermSum())


@pytest.mark.parametrize("policy", [TruncationPolicy(weight_cutoff=3), TruncationPolicy(coeff_cutoff=0.4),
                                    TruncationPolicy(weight_cutoff=5, coeff_cutoff=0.2)])
def test_truncate_matches_policy(policy):
    ts = _pauli_sum(12, 300, 7)
    keep = {t.key: c for t, c in ts.items() if not policy.should_truncate(t.weight, abs(c))}
    ts.truncate(policy)
    assert ts._terms == keep


def test_damping_matches_per_term():
    class Damp:
        def damping_factor(self, weight, active):
            return 0.9 ** weight + 0.01 * active

    ts = _majorana_sum(40, 150, 9)
    want = {t.key: c * Damp().damping_factor(jw_weight(t.modes, 20), 3) for t, c in ts.items()}
    ts.apply_damping(Damp(), 3)
    assert ts._terms == want


def test_key_array_widths():
    ts = _pauli_sum(20, 10, 4)
    k = ts.key_array()
    assert k.shape == (10, 1)
    assert np.array_equal(ts.key_array(3)[:, 0], k[:, 0]) and not ts.key_array(3)[:, 1:].any()
    assert [int(v) for v in k[:, 0]] == [t.key for t, _ in ts.items()]
    np.testing.assert_array_equal(ts.coeff_array(), [c for _, c in ts.items()])


@pytest.mark.parametrize("cls,unit", [(PauliTermSum, lambda: PauliString(1, 0, 70)),
                                      (MajoranaTermSum, lambda: MajoranaMonomial(3, 130))])
def test_empty_sums_save_and_load(tmp_path, cls, unit):
    for ts in (cls(), cls._from_arrays(np.zeros((0, 3), dtype=np.uint64), np.zeros(0), 130)):
        f = tmp_path / "e.gz"
        ts.save(str(f))
        back = cls.from_file(str(f))
        assert len(back) == 0 and back._terms == {}
        assert len(ts.weights()) == 0
        assert ts.key_array().shape[0] == 0
    ts = cls()
    ts.add(unit(), 1.0)
    ts.truncate(TruncationPolicy(coeff_cutoff=2.0))
    assert len(ts) == 0
    ts.save(str(tmp_path / "e2.gz"))
    assert len(cls.from_file(str(tmp_path / "e2.gz"))) == 0
