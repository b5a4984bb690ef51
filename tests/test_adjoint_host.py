"""Host-side pieces of the adjoint path (no GPU): the Heisenberg-order map
of rotations back to circuit order, the checkpoint segmentation, and the
parameter-shift rule the GPU gradient tests use as their exact reference,
pinned here on dense matrices against central differences."""

import math

import numpy as np
import pytest

from oracle import oracle as orc
from propaq_amd import adjoint
from tests.parity import random_case
from tests.test_oracle_dense import dense_heisenberg
from tests.test_oracle_pins import jw_gammas, majorana_dense, pauli_dense


def test_heisenberg_order_maps_back_to_circuit_order():
    rng = np.random.default_rng(0)
    layers = [[(int(rng.integers(1, 1 << 20)), float(i * 10 + j)) for j in range(int(rng.integers(0, 5)))]
              for i in range(6)]
    heis, order = adjoint.heisenberg_layers(layers, 1)
    flat_circuit = [a for layer in layers for _, a in layer]
    flat_heis = [a for _, angles in heis for a in angles]
    # the j-th applied gate is circuit gate order[j]: last layer first, its gates last first
    assert [flat_circuit[i] for i in order] == flat_heis
    assert sorted(order.tolist()) == list(range(len(flat_circuit)))
    g = np.arange(len(order), dtype=np.float64)
    back = adjoint.to_circuit_order(g, order)
    assert all(back[order[j]] == g[j] for j in range(len(order)))


@pytest.mark.parametrize("L,k,expect", [(7, 2, [2, 2, 2, 1]), (4, 4, [4]), (5, 0, [5]), (3, None, [3]), (0, 3, [0])])
def test_checkpoint_segments(L, k, expect):
    heis = [(np.zeros((1, 1), dtype=np.uint64), np.zeros(1)) for _ in range(L)]
    s = adjoint.AdjointSession(None, None, None, heis, None, 0, k)
    assert [len(seg) for seg in s.segments] == expect
    assert s.n_gates == L


def shifted(layers, j, d):
    flat = [list(layer) for layer in layers]
    pos = 0
    for layer in flat:
        for i, (g, a) in enumerate(layer):
            if pos == j:
                layer[i] = (g, a + d)
            pos += 1
    return flat


@pytest.mark.parametrize("kind", [orc.PAULI, orc.MAJORANA])
@pytest.mark.parametrize("seed", range(3))
def test_parameter_shift_is_the_exact_derivative(kind, seed):
    """(f(t + pi/2) - f(t - pi/2)) / 2 equals the derivative of the dense
    expectation for exp(-i t G / 2) rotations, noisy or not."""
    n = 3
    obs, layers, diag = random_case(300 + seed, kind, n, n_layers=3, gates_per_layer=3)
    diag &= (1 << n) - 1
    if kind == orc.PAULI:
        mats = lambda g: pauli_dense(g, n)
        O = sum(c * pauli_dense(k, n) for k, c in obs)
        damping = 0.05
    else:
        gam = jw_gammas(n)
        mats = lambda g: majorana_dense(g, gam)
        O = sum(c * majorana_dense(k, gam) for k, c in obs)
        damping = 0.05
    f = lambda ls: dense_heisenberg(O, ls, mats, damping, n, diag)
    n_rot = sum(len(l) for l in layers)
    h = 1e-5
    for j in range(n_rot):
        ps = (f(shifted(layers, j, math.pi / 2)) - f(shifted(layers, j, -math.pi / 2))) / 2
        fd = (f(shifted(layers, j, h)) - f(shifted(layers, j, -h))) / (2 * h)
        assert abs(ps - fd) < 1e-8, (j, ps, fd)
