"""Whole-run pins of the oracle against dense-matrix evolution.

Without truncation the Heisenberg propagation is exact, so
<psi0| U^dag O U |psi0> from dense matrices fixes the expectation value for
any circuit (the reference checks the same against qiskit Statevector /
DensityMatrix, tests/propagator/test_pauli_random.py:36 and
tests/propagator/test_cross_propagator.py:74-151).  Uniform noise exp(-g w)
per layer is applied in the Pauli basis of the dense operator, at the start of
every layer in Heisenberg order, as the run driver does
(crates/pauli/src/engine.rs:119-160).

Noisy runs switch Clifford deferral off: the reference damps the stored,
pre-frame key (termsum.rs:321 reads ``self.store.row(i)``), so with a deferred
weight-changing Clifford and no weight cutoff its noise acts on the frame's
pre-image weight, which a dense channel does not reproduce.  The oracle and
the device keep that behaviour; the dense check covers the physical one.
"""

import math

import numpy as np
import pytest

from oracle import oracle as orc
from tests.parity import random_case
from tests.test_oracle_pins import jw_gammas, majorana_dense, pauli_dense


def rotation(G):
    n = G.shape[0]
    return lambda th: math.cos(th / 2) * np.eye(n) - 1j * math.sin(th / 2) * G


def pauli_basis(n):
    return [(k, pauli_dense(k, n)) for k in range(4 ** n)]


def damp(O, n, damping, basis):
    """O -> sum_P exp(-damping * weight(P)) tr(P O)/2^n P."""
    out = np.zeros_like(O)
    for k, P in basis:
        c = np.trace(P @ O) / 2 ** n
        if abs(c) < 1e-15:
            continue
        w = sum(1 for q in range(n) if (k >> (2 * q)) & 3)
        out += math.exp(-damping * w) * c * P
    return out


def dense_heisenberg(O, layers, mats, damping, n, psi0_bits):
    basis = pauli_basis(n) if damping is not None else None
    for layer in reversed(layers):
        if damping is not None:
            O = damp(O, n, damping, basis)
        for g, th in reversed(layer):
            U = rotation(mats(g))(th)
            O = U.conj().T @ O @ U
    psi = np.zeros(2 ** n)
    psi[psi0_bits] = 1.0
    return float(np.real(psi @ O @ psi))


@pytest.mark.parametrize("seed", range(12))
@pytest.mark.parametrize("noisy", [False, True])
def test_pauli_runs_match_dense(seed, noisy):
    n = 3
    obs, layers, diag = random_case(seed, orc.PAULI, n, n_layers=4, gates_per_layer=4)
    diag &= (1 << n) - 1
    damping = 0.07 if noisy else None
    table = (lambda li, nl: [math.exp(-damping * w) for w in range(n + 1)]) if noisy else None
    r = orc.run_oracle(orc.PAULI, n, obs, layers, noise_table=table, diag_state=diag, defer=not noisy)
    O = sum(c * pauli_dense(k, n) for k, c in obs)
    want = dense_heisenberg(O, layers, lambda g: pauli_dense(g, n), damping, n, diag)
    assert abs(r.expectation_value - want) < 1e-10


@pytest.mark.parametrize("seed", range(12))
@pytest.mark.parametrize("noisy", [False, True])
def test_majorana_runs_match_dense(seed, noisy):
    n = 3
    gam = jw_gammas(n)
    obs, layers, diag = random_case(50 + seed, orc.MAJORANA, n, n_layers=4, gates_per_layer=4)
    diag &= (1 << n) - 1
    damping = 0.07 if noisy else None
    table = (lambda li, nl: [math.exp(-damping * w) for w in range(n + 1)]) if noisy else None
    r = orc.run_oracle(orc.MAJORANA, n, obs, layers, noise_table=table, diag_state=diag, defer=not noisy)
    O = sum(c * majorana_dense(k, gam) for k, c in obs)
    want = dense_heisenberg(O, layers, lambda g: majorana_dense(g, gam), damping, n, diag)
    assert abs(r.expectation_value - want) < 1e-10


@pytest.mark.parametrize("seed", range(6))
def test_frame_deferral_is_exact(seed):
    """Runs full of Clifford angles with the frame on and off agree, and agree
    with dense evolution (the frame is only a different bookkeeping)."""
    n = 3
    obs, layers, diag = random_case(80 + seed, orc.PAULI, n, n_layers=4, gates_per_layer=5, clifford_frac=0.6)
    diag &= (1 << n) - 1
    a = orc.run_oracle(orc.PAULI, n, obs, layers, diag_state=diag, defer=True, want_terms=True)
    b = orc.run_oracle(orc.PAULI, n, obs, layers, diag_state=diag, defer=False, want_terms=True)
    O = sum(c * pauli_dense(k, n) for k, c in obs)
    want = dense_heisenberg(O, layers, lambda g: pauli_dense(g, n), None, n, diag)
    assert abs(a.expectation_value - want) < 1e-10 and abs(b.expectation_value - want) < 1e-10
    ka = {k: v for k, v in a.terms.items() if abs(v) > 1e-12}
    kb = {k: v for k, v in b.terms.items() if abs(v) > 1e-12}
    assert set(ka) == set(kb)
    for k in ka:
        assert abs(ka[k] - kb[k]) < 1e-12


def test_norm_invariant_without_truncation():  # tests/propagator/test_pauli_propagator.py:90
    obs, layers, _ = random_case(7, orc.PAULI, 10, n_layers=6, gates_per_layer=8, clifford_frac=0.0)
    r = orc.run_oracle(orc.PAULI, 10, obs, layers, want_terms=True)
    assert abs(sum(c * c for c in r.terms.values()) - sum(c * c for _, c in obs)) < 1e-10


def test_loschmidt_echo():  # tests/propagator/test_pauli_loschmidt.py:15
    obs, layers, _ = random_case(9, orc.PAULI, 8, n_layers=5, gates_per_layer=6, clifford_frac=0.0)
    inv = [[(g, -a) for g, a in reversed(layer)] for layer in reversed(layers)]
    r = orc.run_oracle(orc.PAULI, 8, obs, layers + inv, want_terms=True)
    got = {k: v for k, v in r.terms.items() if abs(v) > 1e-12}
    want = {}
    for k, c in obs:
        want[k] = want.get(k, 0.0) + c
    assert set(got) == set(want)
    for k, c in want.items():
        assert abs(got[k] - c) < 1e-10
