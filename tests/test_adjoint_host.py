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


class FakeEngine:
    """Stands in for DeviceEngine in the segmentation logic: every gate
    multiplies the row count by 1.5, a recorded gate adds 16 bytes per input
    row to the tape, and every call is logged."""

    def __init__(self):
        self.n, self.rec, self.tape, self.gates_on_tape, self.log = 0, False, 0, 0, []
        self.ckpt = {}
        self.budget, self.swept = None, 0

    def set_differentiable(self, on):
        pass

    def counters(self):
        return {"term_gate_updates": 0}

    def seed(self, keys, coeffs):
        self.n = 4

    def __len__(self):
        return self.n

    def apply_gates(self, gens, angles):
        out = []
        for a in angles:
            if self.rec:
                self.tape += 16 * self.n
                self.gates_on_tape += 1
            self.log.append(("gate", float(a), self.rec))
            self.n = int(self.n * 1.5) + 1
            out.append(self.n)
        return np.array(out)

    def tape_info(self):
        return self.gates_on_tape, self.tape

    def tape_begin(self):
        self.rec, self.tape, self.gates_on_tape = True, 0, 0
        self.tape_start_n = self.n

    def noise(self):
        if self.rec:
            self.tape += 12 * self.n
        self.n = self.n * 2 + 3

    def tape_end(self):
        self.rec = False

    def checkpoint_save(self):
        i = len(self.ckpt) + 1
        self.ckpt[i] = self.n
        return i

    def checkpoint_restore(self, i):
        self.n = self.ckpt[i]

    def checkpoint_free(self, i):
        del self.ckpt[i]

    def expectation(self, diag):
        return 0.0

    def synchronize(self):
        pass

    def adjoint_seed(self, diag):
        self.lam_n = self.n

    def adjoint_backward(self, n_gates):
        assert n_gates == self.gates_on_tape
        assert self.n == self.lam_n, "the tape does not end at the adjoint's rows"
        self.lam_n = self.tape_start_n
        # a segment stays within the budget unless it holds one gate at most
        # (with its layer's noise pass)
        assert self.tape <= self.budget or n_gates <= 1, (self.tape, n_gates)
        self.swept += 1
        self.tape, self.gates_on_tape = 0, 0
        return np.zeros(n_gates)


@pytest.mark.parametrize("every,budget", [(2, 1 << 40), (0, 1 << 40), (3, 5000), (1, 1000), (4, 1), (0, 3000),
                                          (0, 20000), (2, 100000)])
def test_segments_cover_every_gate_once_within_budget(every, budget):
    sizes = [3, 0, 5, 2, 4, 1]
    heis = [(np.zeros((m, 1), dtype=np.uint64), np.arange(m, dtype=np.float64) + 10 * li)
            for li, m in enumerate(sizes)]
    eng = FakeEngine()
    eng.budget = budget
    noise_calls = []
    step = lambda e, li, nl: (noise_calls.append(li), e.log.append(("noise", li, e.rec)), e.noise())
    s = adjoint.AdjointSession(eng, None, None, heis, step, 0, every, budget)
    s.forward()
    assert noise_calls == list(range(len(sizes)))
    fwd = [x for x in eng.log if x[0] == "gate"]
    assert [a for _, a, _ in fwd] == [float(a) for _, ang in heis for a in ang]
    assert all(r for _, _, r in fwd)  # the forward run records throughout
    # segments tile the gate sequence
    flat = [(li, g) for li, m in enumerate(sizes) for g in range(m)]
    covered = []
    for (l0, g0), (l1, g1) in s.segments:
        covered += [p for p in flat if (l0, g0) <= p < (l1, g1)]
    assert covered == flat
    if every:
        for (l0, _), (l1, g1) in s.segments:
            assert l1 - l0 <= every or (l1 - l0 == every and g1 == 0)
    # the sweep replays every earlier segment exactly as the forward run did
    eng.log.clear()
    g = s.backward()
    assert len(g) == sum(sizes)
    replayed = [x for x in eng.log if x[0] == "gate"]
    expect = []
    for sg in reversed(s.segments[:-1]):
        (l0, g0), (l1, g1) = sg
        expect += [float(heis[li][1][gi]) for li, gi in flat if (l0, g0) <= (li, gi) < (l1, g1)]
    assert [a for _, a, _ in replayed] == expect
    replay_noise = [x[1] for x in eng.log if x[0] == "noise"]
    expect_noise = []
    for (l0, g0), (l1, g1) in reversed(s.segments[:-1]):
        expect_noise += [li for li in range(l0, min(l1 + 1, len(sizes)))
                         if (li > l0 or g0 == 0) and (li < l1 or g1 > 0)]
    assert replay_noise == expect_noise
    assert not eng.ckpt  # every checkpoint freed
    assert eng.swept == len(s.segments)


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
