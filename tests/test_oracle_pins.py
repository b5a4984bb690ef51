"""Pins of the FP64 oracle against the reference's own known-answer tests.

Each test restates one reference unit test (cited by file:line) with the
same inputs and expected values.  Where the reference test only compares two
of its own implementations, the expected value is fixed here by mathematics
instead (dense matrices, closed forms), so a dropped term, a wrong sign or a
transposed operand in the oracle fails a check.
"""

import math

import numpy as np
import pytest

from oracle import oracle as orc

TOY, PAULI, MAJ = orc.TOY, orc.PAULI, orc.MAJORANA


def mono(*bits):
    k = 0
    for b in bits:
        k |= 1 << b
    return k


def toy(n_units=8, **cut):
    e = orc.OracleEngine(TOY, n_units, 1)
    e.set_cutoff(cut.get("max_weight"), cut.get("min_coeff"), cut.get("min_terms"))
    e.set_defer(False)
    return e


# --- crates/core/tests/unit/engine/termsum.rs -----------------------------

def test_add_folds_duplicate_key():  # termsum.rs:53
    e = toy()
    e.add(mono(0), 1.0)
    e.add(mono(1), 2.0)
    e.add(mono(0), 3.0)
    assert len(e) == 2
    v = e.terms()
    assert v[0b01] == 4.0 and v[0b10] == 2.0


def test_commuting_term_untouched():  # termsum.rs:65
    e = toy()
    e.add(mono(1), 5.0)
    assert e.apply_rotation(mono(0), 0.7) == 0
    assert e.terms() == {0b10: 5.0}


def test_anticommuting_splits_cos_sin():  # termsum.rs:78
    e = toy()
    e.add(mono(0), 2.0)
    assert e.apply_rotation(mono(0), 0.3) == 1
    v = e.terms()
    assert abs(v[0b01] - 2.0 * math.cos(0.3)) < 1e-12
    assert abs(v[0b00] - -(2.0 * math.sin(0.3))) < 1e-12


def test_child_on_existing_row_accumulates():  # termsum.rs:100
    e = toy()
    e.add(mono(0), 2.0)
    e.add(mono(1), 3.0)
    assert e.apply_rotation(mono(0, 1), 0.3) == 0
    assert len(e) == 2


def test_sine_from_pre_rotation_coefficient():  # termsum.rs:117
    a, c0, c1 = 0.4, 2.0, 5.0
    e = toy()
    e.add(mono(0), c0)
    e.add(mono(1), c1)
    e.apply_rotation(mono(0, 1), a)
    v = e.terms()
    s, c = math.sin(a), math.cos(a)
    assert abs(v[0b01] - (c0 * c - c1 * s)) < 1e-12
    assert abs(v[0b10] - (c1 * c - c0 * s)) < 1e-12


def test_phase_only_rotation_scales():  # termsum.rs:143
    e = toy()
    e.add(mono(0), 2.0)
    e.add(mono(1), 3.0)
    assert e.apply_rotation(mono(0), math.pi) == 0
    v = e.terms()
    assert abs(v[0b01] + 2.0) < 1e-12 and v[0b10] == 3.0


def test_weight_cutoff_suppresses_child():  # termsum.rs:161
    e = toy(max_weight=2)
    e.add(mono(0, 1), 1.0)
    assert e.apply_rotation(mono(2), 0.3) == 0
    assert len(e) == 1


def test_coefficient_cutoff_suppresses_child():  # termsum.rs:176
    e = toy(min_coeff=1e-6)
    e.add(mono(0), 1e-12)
    assert e.apply_rotation(mono(0), 0.3) == 0
    assert len(e) == 1


def test_pair_rescue():  # termsum.rs:191
    a, big, small = 0.3, 1.0, 1e-9
    e = toy(min_coeff=1e-6)
    e.add(mono(0), big)
    e.add(mono(1), small)
    e.apply_rotation(mono(0, 1), a)
    v = e.terms()
    s, c = math.sin(a), math.cos(a)
    assert len(e) == 2
    assert abs(v[0b01] - (big * c - small * s)) < 1e-18
    assert abs(v[0b10] - (small * c - big * s)) < 1e-15


def test_term_floor_suppresses_lossy_predicates():  # termsum.rs:219
    tiny = 1e-9
    e = toy(min_coeff=1e-6, min_terms=1000)
    e.add(mono(0), tiny)
    e.apply_rotation(mono(0), 0.3)
    assert len(e) == 2
    e = toy(min_coeff=1e-6)
    e.add(mono(0), tiny)
    e.apply_rotation(mono(0), 0.3)
    assert len(e) == 1


def test_pair_rescue_needs_one_earned_half():  # termsum.rs:285
    e = toy(min_coeff=1e-6)
    e.add(mono(0), 1e-9)
    e.add(mono(1), 2e-9)
    e.apply_rotation(mono(0, 1), 0.3)
    v = e.terms()
    c = math.cos(0.3)
    assert len(e) == 2
    assert abs(v[0b01] - 1e-9 * c) < 1e-24 and abs(v[0b10] - 2e-9 * c) < 1e-24


def test_pair_rescue_needs_both_halves():  # termsum.rs:305
    e = toy(min_coeff=1e-6)
    e.add(mono(0), 1e-9)
    assert e.apply_rotation(mono(0), 0.3) == 0
    assert len(e) == 1


def test_orbit_dedup():  # termsum.rs:320
    e = toy()
    e.add(mono(0), 1.0)
    for _ in range(12):
        e.apply_rotation(mono(0), 0.3)
    assert len(e) == 2


def test_expectation_coeff_times_trace():  # termsum.rs:332
    e = toy()
    e.add(mono(0), 2.0)
    e.add(mono(1), 3.0)
    assert abs(e.expectation(0b01) - 1.0) < 1e-12


def test_rows_keep_their_index():  # termsum.rs:341
    e = toy()
    e.add(mono(0), 1.0)
    for _ in range(8):
        e.apply_rotation(mono(1), 0.3)
    keys, _ = e.export(frame=False)
    assert int(keys[0][0]) == 0b01


def test_weight_only_factors_match_table_bit_for_bit():  # termsum.rs:414
    damping = 0.37
    keys = [mono(0), mono(1), mono(0, 1), mono(2, 3, 4)]
    by_table, by_factor = toy(), toy()
    for i, k in enumerate(keys):
        by_table.add(k, 1.0 + i)
        by_factor.add(k, 1.0 + i)
    by_table.scale_by_table([math.exp(-damping * w) for w in range(9)])
    f = np.array([math.exp(-damping * bin(k).count("1")) for k in keys])
    by_factor.scale_by_factors(f)
    assert by_table.terms() == by_factor.terms()


def test_reclaim_by_flags():  # termsum.rs:474
    e = toy()
    for k in [mono(0), mono(1), mono(2), mono(0, 2)]:
        e.add(k, 1.0)
    keep = np.array([(k & ~0b0101) == 0 for k in [mono(0), mono(1), mono(2), mono(0, 2)]], dtype=np.uint8)
    assert e.reclaim(keep) == 1
    assert set(e.terms()) == {mono(0), mono(2), mono(0, 2)}


# --- crates/core/tests/unit/engine/partitioned_termsum.rs ------------------

def test_reclaim_drops_decayed_terms_and_store_still_branches():  # partitioned_termsum.rs:177
    e = toy(min_coeff=0.1)
    for k in range(24):
        e.add(mono(k % 6, k // 6 + 1), 1.0)
    before = len(e)
    e.scale_by_table([1e-6] * 9)
    assert e.reclaim() == before and len(e) == 0
    e.set_cutoff(None, None, None)
    e.add(mono(0), 1.0)
    e.apply_rotation(mono(0), 0.3)
    assert len(e) == 2


def test_reclaim_keeps_admitted():  # partitioned_termsum.rs:205
    e = toy(min_coeff=0.1)
    e.add(mono(0), 1.0)
    e.add(mono(1), 1e-9)
    e.add(mono(2), 0.5)
    assert e.reclaim() == 1 and len(e) == 2


def test_reclaim_without_cutoff_is_noop():  # partitioned_termsum.rs:220
    e = toy()
    e.add(mono(0), 1e-30)
    assert e.reclaim() == 0 and len(e) == 1


def test_empty_operator_rotation_is_noop():  # partitioned_termsum.rs:319
    e = toy()
    assert e.apply_rotation(mono(0), 0.3) == 0 and len(e) == 0


# --- crates/core/tests/unit/policy/truncators.rs:94,105 --------------------

def test_cutoff_boundaries_keep():
    e = orc.OracleEngine(PAULI, 8)
    e.set_cutoff(3, 0.5)
    three = (1 << 1) | (1 << 3) | (1 << 5)  # Z0 Z1 Z2, weight 3
    four = three | (1 << 7)
    assert e.admits_initial(three, 0.5)
    assert not e.admits_initial(four, 0.5)
    assert not e.admits_initial(three, 0.4999)


# --- crates/core/tests/unit/algebra/strings.rs ------------------------------

def test_support_counts_units():  # strings.rs:293,301
    assert orc.weight(PAULI, mono(0, 1, 6), 32, 1) == 2
    assert orc.weight(PAULI, mono(62, 63, 64), 64, 2) == 2


# --- crates/pauli/tests/unit/string.rs --------------------------------------

PAULIS = {
    "I": np.eye(2, dtype=complex),
    "X": np.array([[0, 1], [1, 0]], dtype=complex),
    "Y": np.array([[0, -1j], [1j, 0]], dtype=complex),
    "Z": np.array([[1, 0], [0, -1]], dtype=complex),
}


def pauli_dense(key, n):
    m = np.array([[1.0 + 0j]])
    for q in reversed(range(n)):
        xb, zb = (key >> (2 * q)) & 1, (key >> (2 * q + 1)) & 1
        p = "Y" if xb and zb else ("X" if xb else ("Z" if zb else "I"))
        m = np.kron(m, PAULIS[p])
    return m


def test_x_times_z_is_minus_i_y():  # string.rs:75
    key, e = orc.product(PAULI, mono(0), mono(1), 1)
    assert key == mono(0, 1) and e == 3


def test_pauli_product_phases_exhaustive_two_qubits():  # string.rs:196 (dense matrices fix the phase)
    for a in range(16):
        for b in range(16):
            key, e = orc.product(PAULI, a, b, 2)
            assert key == a ^ b
            lhs = pauli_dense(a, 2) @ pauli_dense(b, 2)
            assert np.allclose(lhs, (1j ** e) * pauli_dense(key, 2))
            anti = np.allclose(lhs, -(pauli_dense(b, 2) @ pauli_dense(a, 2)))
            assert orc.anticommutes(PAULI, a, b, 2) == anti


@pytest.mark.parametrize("key,state,want", [
    (0, 0, 1.0), (mono(0), 0, 0.0), (mono(1), 0b00, 1.0), (mono(1), 0b01, -1.0),
    (mono(1, 3), 0b00, 1.0), (mono(1, 3), 0b01, -1.0), (mono(1, 3), 0b10, -1.0), (mono(1, 3), 0b11, 1.0),
])
def test_pauli_trace(key, state, want):  # string.rs:86-120
    assert orc.trace(PAULI, key, 4, state) == want


# --- crates/majorana/tests/unit/monomial.rs ---------------------------------

@pytest.mark.parametrize("bits,n_modes,want", [
    (0, 8, 0), (0b01, 8, 1), (0b11, 8, 1), (0b01010101, 8, 4), (0b01, 128, 1), (1 << 64, 128, 33),
])
def test_majorana_jw_weight(bits, n_modes, want):  # monomial.rs:77-105
    assert orc.weight(MAJ, bits, n_modes // 2) == want


@pytest.mark.parametrize("bits,state,want", [
    (0, 0, 1.0), (0, 0b1111, 1.0), (0b01, 0, 0.0), (0b01, 1, 0.0), (0b11, 0, -1.0), (0b11, 1, 1.0),
    (0b1111, 0b00, -1.0), (0b1111, 0b01, 1.0), (0b1111, 0b10, 1.0), (0b1111, 0b11, -1.0),
])
def test_majorana_trace(bits, state, want):  # monomial.rs:108-138
    assert orc.trace(MAJ, bits, 4, state) == want


@pytest.mark.parametrize("a,b,want_e", [(0, 0b0011, 0), (0b0011, 0, 0), (0b0111, 0b0111, 0), (0b0011, 0b1100, 2)])
def test_majorana_matmul_phases(a, b, want_e):  # monomial.rs:156-200
    key, e = orc.product(MAJ, a, b, 4)
    assert key == a ^ b and e == want_e


@pytest.mark.parametrize("a,b,anti", [(0b0011, 0b0011, False), (0b0011, 0b1100, False), (0b0011, 0b0110, True),
                                      (0b0001, 0b0010, True)])
def test_majorana_commutation(a, b, anti):  # monomial.rs:203-226
    assert orc.anticommutes(MAJ, a, b, 4) == anti


def jw_gammas(n_sites):
    """gamma_2k = Z..Z X_k, gamma_2k+1 = Z..Z Y_k as dense matrices."""
    g = []
    for k in range(n_sites):
        for p in ("X", "Y"):
            m = np.array([[1.0 + 0j]])
            for q in reversed(range(n_sites)):
                m = np.kron(m, PAULIS["Z"] if q < k else (PAULIS[p] if q == k else PAULIS["I"]))
            g.append(m)
    return g


def majorana_dense(bits, gam):
    """i^h(L) gamma_a1 ... gamma_aL with ascending modes (Hermitian form)."""
    modes = [i for i in range(len(gam)) if (bits >> i) & 1]
    m = np.eye(gam[0].shape[0], dtype=complex)
    for i in modes:
        m = m @ gam[i]
    L = len(modes)
    return (1j ** (0 if L % 4 in (0, 1) else 1)) * m


def test_majorana_products_against_dense_jw():
    n_sites = 3
    gam = jw_gammas(n_sites)
    rng = np.random.default_rng(7)
    for _ in range(200):
        a, b = int(rng.integers(0, 64)), int(rng.integers(0, 64))
        key, e = orc.product(MAJ, a, b, n_sites)
        A, B = majorana_dense(a, gam), majorana_dense(b, gam)
        assert key == a ^ b
        assert np.allclose(A @ B, (1j ** e) * majorana_dense(key, gam))
        assert orc.anticommutes(MAJ, a, b, n_sites) == np.allclose(A @ B, -(B @ A))


def test_majorana_weight_is_weight_of_jw_image():  # monomial.rs:256,275 pinned by the JW map itself
    n_sites = 4
    gam = jw_gammas(n_sites)
    for bits in range(0, 1 << (2 * n_sites), 7):
        M = majorana_dense(bits, gam)
        # weight = number of qubits where the JW image acts non-trivially
        w = 0
        for q in range(n_sites):
            # partial trace test: M commutes with all single-qubit Paulis on q iff identity there
            acts = False
            for p in ("X", "Z"):
                m = np.array([[1.0 + 0j]])
                for qq in reversed(range(n_sites)):
                    m = np.kron(m, PAULIS[p] if qq == q else PAULIS["I"])
                if not np.allclose(M @ m, m @ M):
                    acts = True
            w += acts
        assert orc.weight(MAJ, bits, n_sites) == w, bin(bits)


# --- crates/core/tests/unit/algebra/tableau.rs (W = 2, 8 units) --------------

def X(q):
    return 1 << (2 * q)


def Z(q):
    return 1 << (2 * q + 1)


QUARTER = math.pi / 2


def frame(*steps):
    e = orc.OracleEngine(PAULI, 8, 2)
    for g in steps:
        ok, _ = e.frame_compose_rotation(g, QUARTER)
        assert ok
    return e


def test_fresh_tableau_is_identity():  # tableau.rs:80
    e = orc.OracleEngine(PAULI, 8, 2)
    assert e.frame_is_identity
    for m in [X(0), Z(1), X(0) | Z(2)]:
        assert e.frame_conjugate(m) == (m, 1.0)


def test_single_qubit_quarter_turn_maps_x_to_y():  # tableau.rs:90
    e = frame(Z(0))
    assert e.frame_conjugate(X(0))[0] == X(0) | Z(0)
    assert e.frame_conjugate(Z(0))[0] == Z(0)


def test_quarter_turn_signs_against_dense_conjugation():  # tableau.rs:90, sign fixed by matrices
    for gen in [Z(0), Z(0) | Z(1), X(1), X(0) | X(1)]:
        e = frame(gen)
        U = np.cos(QUARTER / 2) * np.eye(4) - 1j * np.sin(QUARTER / 2) * pauli_dense(gen, 2)
        for m in range(1, 16):
            img, s = e.frame_conjugate(m)
            # readout map P -> U^dag P U, the Heisenberg image of the rotation
            assert np.allclose(U.conj().T @ pauli_dense(m, 2) @ U, s * pauli_dense(img, 2))


def test_two_qubit_quarter_turn_changes_weight():  # tableau.rs:107
    e = orc.OracleEngine(PAULI, 8, 2)
    ok, cw = e.frame_compose_rotation(Z(0) | Z(1), QUARTER)
    assert ok and cw
    assert orc.weight(PAULI, e.frame_conjugate(X(0))[0], 8, 2) == 2
    ok, cw = orc.OracleEngine(PAULI, 8, 2).frame_compose_rotation(Z(0), QUARTER)
    assert ok and not cw


def test_conjugation_injective_and_signs_unit():  # tableau.rs:144,158
    e = frame(Z(0) | Z(1))
    seen = set()
    for bits in range(1024):
        img, s = e.frame_conjugate(bits)
        assert img not in seen and abs(s) == 1.0
        seen.add(img)


def test_four_quarter_turns_are_identity():  # tableau.rs:169
    e = frame(*([Z(0) | Z(1)] * 4))
    for bits in range(512):
        assert e.frame_conjugate(bits) == (bits, 1.0)


def test_compose_matches_sequential():  # tableau.rs:186
    a, b, ab = frame(Z(0)), frame(Z(0) | Z(1)), frame(Z(0), Z(0) | Z(1))
    for bits in range(512):
        once, s1 = ab.frame_conjugate(bits)
        mid, sa = a.frame_conjugate(bits)
        twice, sb = b.frame_conjugate(mid)
        assert once == twice and s1 == sa * sb


def test_two_directions_are_inverses():  # tableau.rs:205
    acc = frame(Z(0), Z(0) | Z(1), X(2), Z(1) | Z(2))
    for bits in range(1024):
        fwd, s1 = acc.frame_conjugate(bits)
        back, s2 = acc.frame_conjugate(fwd, generator_direction=True)
        assert back == bits and s1 * s2 == 1.0


def test_non_clifford_angle_rejected():  # tableau.rs:223
    e = orc.OracleEngine(PAULI, 8, 2)
    assert e.frame_compose_rotation(Z(0), 0.37) == (False, False)
