"""Adjoint differentiation of the propagated expectation value.

The reference has no gradient API; its nearest equivalent is parameter shift,
two forward runs at theta +- pi/2 per angle through
``PauliPropagator.expectation_value`` (``crates/pauli/src/propagator.rs:270``).
Here d<O>/dtheta of every rotation exp(-i theta G / 2) comes from one forward
run and one reverse sweep over a gather tape recorded by the device kernels
(``pq_tape_begin`` .. ``pq_adjoint_backward`` in ``include/propaq_amd.h``):

* forward: the layers run in Heisenberg order as in ``run_at_width``
  (``crates/pauli/src/engine.rs:74-238``) with recording on; every segment
  (``checkpoint_every`` layers, cut shorter where its tape would exceed a
  byte budget) starts with a device checkpoint of the rows and a fresh tape;
* backward: the adjoint starts as the readout weight of every final row and
  sweeps the last segment's tape; each earlier segment is recomputed from its
  checkpoint with recording on, then swept.  Recomputation is exact (the
  device path is deterministic), so the replayed rows match the rows the
  adjoint is defined on.

Every truncation decision of the forward run (weight and coefficient
refusals, pending branches, pair-rule claims, reclaim) is held fixed: the
gradient is that of the computed, truncated expectation value.  The engine
runs in differentiable mode (``pq_set_differentiable``): Clifford rotations
are applied as rotations instead of being folded into the frame, and gates at
|sin theta| < 1e-9 form their (near-zero) branches instead of taking the
phase-only shortcut, so every gate's derivative is present at any angle.

``autograd_function()`` wraps a session as a ``torch.autograd.Function``;
``expectation`` is the user-facing call.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

from ._native import DeviceEngine
from .runner import gates_to_arrays, keys_to_array

# noise_step(engine, heisenberg_layer_index, n_layers) runs the layer-start
# noise / reclaim pass on the engine.
NoiseStep = Callable[[DeviceEngine, int, int], None]


@dataclass
class GradientRun:
    expectation_value: float
    gradient: np.ndarray  # d<O>/dtheta per rotation, circuit order (layers, then gates)
    n_terms: List[int] = field(default_factory=list)  # live count after each gate, Heisenberg order
    forward_s: float = 0.0
    backward_s: float = 0.0  # recomputation included
    term_gate_updates: int = 0  # of the forward run
    tape_bytes_peak: int = 0
    checkpoints: int = 0


class AdjointSession:
    """One forward run kept on the device for a later reverse sweep.

    ``heis_layers`` holds (generators uint64[n, W], angles float64[n]) per
    layer in Heisenberg order (the order the gates are applied).

    The forward run records as it goes.  A new segment starts, with a device
    checkpoint of the rows and a fresh tape, every ``checkpoint_every``
    layers, and also before any gate whose tape could push the current
    segment past ``tape_budget`` bytes (a gate's tape is 16 bytes per input
    row: two u32 slot maps and the FP64 coefficients; the bound assumes the
    row count may double at every gate).  At the end the tape holds
    the last segment; the sweep replays the earlier ones from their
    checkpoints, last first."""

    def __init__(self, eng: DeviceEngine, seed_keys: np.ndarray, seed_coeffs: np.ndarray,
                 heis_layers: Sequence[Tuple[np.ndarray, np.ndarray]], noise_step: Optional[NoiseStep] = None,
                 diag_state: int = 0, checkpoint_every: Optional[int] = 4, tape_budget: int = 64 << 30):
        self.eng = eng
        self.seed_keys, self.seed_coeffs = seed_keys, seed_coeffs
        self.layers = list(heis_layers)
        self.noise_step = noise_step
        self.diag_state = int(diag_state)
        L = len(self.layers)
        self.every = int(checkpoint_every) if checkpoint_every and checkpoint_every > 0 else max(L, 1)
        self.tape_budget = int(tape_budget)
        self.n_gates = sum(len(a) for _, a in self.layers)
        # segments: [start, end) positions (layer, gate); noise runs at a
        # segment's start only when it begins a layer
        self.segments: List[Tuple[Tuple[int, int], Tuple[int, int]]] = []
        self._ckpts: List[int] = []
        self._done_forward = False
        self.tape_bytes_peak = 0
        self.n_terms: List[int] = []
        self.forward_s = self.backward_s = 0.0
        self.term_gate_updates = 0

    # -- the engine calls that differ between unbatched and batched sessions
    def _seed(self):
        self.eng.seed(self.seed_keys, self.seed_coeffs)

    def _apply(self, gens, angles):
        return self.eng.apply_gates(gens, angles)

    def _value(self):
        return self.eng.expectation(self.diag_state)

    def _adjoint_seed(self):
        self.eng.adjoint_seed(self.diag_state)

    def _adjoint_backward(self, n_gates: int) -> np.ndarray:
        return self.eng.adjoint_backward(n_gates)

    def _gates(self, li: int, g0: int, g1: int, collect: bool):
        gens, angles = self.layers[li]
        if g1 > g0:
            counts = self._apply(gens[g0:g1], angles[g0:g1])
            if collect:
                self.n_terms.extend(int(v) for v in counts)

    def _replay(self, seg):
        (l0, g0), (l1, g1) = seg
        L = len(self.layers)
        for li in range(l0, min(l1 + 1, L)):
            a = g0 if li == l0 else 0
            b = g1 if li == l1 else len(self.layers[li][1])
            if a == 0 and self.noise_step is not None and (li < l1 or b > 0):
                self.noise_step(self.eng, li, L)
            self._gates(li, a, b, False)

    def _open_segment(self, pos):
        self._ckpts.append(self.eng.checkpoint_save())
        self.eng.tape_begin()
        self._seg_start = pos

    def _close_segment(self, pos):
        self.tape_bytes_peak = max(self.tape_bytes_peak, self.eng.tape_info()[1])
        self.segments.append((self._seg_start, pos))

    def forward(self) -> float:
        eng = self.eng
        eng.set_differentiable(True)
        t0 = time.perf_counter()
        c0 = eng.counters()["term_gate_updates"]
        self._seed()
        L = len(self.layers)
        self._open_segment((0, 0))
        seg_layer = 0
        for li in range(L):
            n_g = len(self.layers[li][1])
            # a segment boundary at a layer start is decided before the
            # layer's noise pass, which belongs to the segment that starts there
            if li > 0:
                n = max(len(eng), 1)
                first = 12 * n + (32 * n if n_g else 0) + 1024  # noise tape + the first gate, doubled
                gates_on_tape, used = eng.tape_info()
                if li - seg_layer >= self.every or (used and used + first > self.tape_budget):
                    self._close_segment((li, 0))
                    self._open_segment((li, 0))
                    seg_layer = li
            if self.noise_step is not None:
                self.noise_step(eng, li, L)
            g = 0
            while g < n_g:
                # gates that fit the budget even if every one doubled the rows
                room = self.tape_budget - eng.tape_info()[1]
                n = max(len(eng), 1)
                fit, cost = 0, 0
                while g + fit < n_g and fit < 64:
                    cost += 16 * n * (2 ** min(fit, 40)) + 512
                    if cost > room:
                        break
                    fit += 1
                if fit == 0:
                    if g > 0:  # start a new segment before this gate
                        self._close_segment((li, g))
                        self._open_segment((li, g))
                        seg_layer = li
                    fit = 1  # a single gate always runs, whatever its size
                self._gates(li, g, g + fit, True)
                g += fit
        eng.tape_end()
        self._close_segment((L, 0))
        # the last segment's tape is live; its checkpoint is not needed
        eng.checkpoint_free(self._ckpts.pop())
        value = self._value()
        eng.synchronize()
        self.forward_s = time.perf_counter() - t0
        self.term_gate_updates = eng.counters()["term_gate_updates"] - c0
        self._done_forward = True
        self.value = value
        return value

    def _seg_gate_count(self, seg) -> int:
        (l0, g0), (l1, g1) = seg
        total = 0
        for li in range(l0, min(l1 + 1, len(self.layers))):
            a = g0 if li == l0 else 0
            b = g1 if li == l1 else len(self.layers[li][1])
            total += max(b - a, 0)
        return total

    def backward(self) -> np.ndarray:
        """d<O>/dtheta per gate in Heisenberg (application) order."""
        if not self._done_forward:
            raise RuntimeError("propaq_amd: backward() before forward()")
        eng = self.eng
        t0 = time.perf_counter()
        self._adjoint_seed()
        last = len(self.segments) - 1
        parts: List[np.ndarray] = [None] * len(self.segments)
        parts[last] = self._adjoint_backward(self._seg_gate_count(self.segments[last]))
        for si in range(last - 1, -1, -1):
            eng.checkpoint_restore(self._ckpts[si])
            eng.tape_begin()
            self._replay(self.segments[si])
            eng.tape_end()
            self.tape_bytes_peak = max(self.tape_bytes_peak, eng.tape_info()[1])
            parts[si] = self._adjoint_backward(self._seg_gate_count(self.segments[si]))
            eng.checkpoint_free(self._ckpts[si])
        self._ckpts = []
        eng.synchronize()
        self.backward_s = time.perf_counter() - t0
        self._done_forward = False  # the tape and checkpoints are consumed
        return np.concatenate(parts) if parts else np.zeros(0)


class BatchAdjointSession(AdjointSession):
    """An ``AdjointSession`` over a batched engine: ``heis_layers`` holds
    (generators uint64[n, W], angles float64[n, n_instances]) per Heisenberg
    step, ``seed_instances`` the instance of every seeded row and
    ``diag_states`` one basis state per instance.  ``forward`` returns the
    expectation value of every instance and ``backward`` d<O_b>/dtheta as
    [gates, n_instances].  Instances that have finished their own layers take
    are frozen (``pq_freeze_instances``) from step ``depth[b]`` on: their
    rows pass through unchanged, so every value is read at the end and every
    instance's gradient equals that of a session of its own.  The frozen set
    of a step is set at its start, also when a segment is replayed."""

    def __init__(self, eng: DeviceEngine, seed_instances: np.ndarray, seed_keys: np.ndarray,
                 seed_coeffs: np.ndarray, heis_layers, diag_states, depth: np.ndarray,
                 checkpoint_every: Optional[int] = 4, tape_budget: int = 64 << 30):
        depth = np.asarray(depth)

        def freeze(e, li, n_layers):
            e.freeze_instances((depth <= li).astype(np.uint8))

        super().__init__(eng, seed_keys, seed_coeffs, heis_layers, freeze, 0, checkpoint_every, tape_budget)
        self.seed_instances = seed_instances
        self.diag_states = diag_states

    def _replay(self, seg):
        # a segment cut between gates starts mid-step: set that step's frozen set
        (l0, g0), _ = seg
        if g0 > 0:
            self.noise_step(self.eng, l0, len(self.layers))
        super()._replay(seg)

    def _seed(self):
        self.eng.seed_batched(self.seed_instances, self.seed_keys, self.seed_coeffs)

    def _apply(self, gens, angles):
        return self.eng.apply_gates_batched(gens, angles)

    def _value(self):
        return self.eng.expectation_batched(self.diag_states)

    def _adjoint_seed(self):
        self.eng.adjoint_seed_batched(self.diag_states)

    def _adjoint_backward(self, n_gates: int) -> np.ndarray:
        return self.eng.adjoint_backward_batched(n_gates)


def heisenberg_layers(layers: Sequence[Sequence[Tuple[int, float]]], W: int):
    """Circuit-order layers of (generator key, angle) -> Heisenberg-order
    arrays, plus the map from Heisenberg gate position to circuit position
    (layers reversed, gates reversed: ``extract_layers``,
    ``crates/pauli/src/engine.rs:44-70``)."""
    starts, pos = [], 0
    for layer in layers:
        starts.append(pos)
        pos += len(layer)
    heis, order = [], []
    for l in range(len(layers) - 1, -1, -1):
        layer = layers[l]
        heis.append(gates_to_arrays(list(reversed(layer)), W))
        order.extend(starts[l] + j for j in range(len(layer) - 1, -1, -1))
    return heis, np.asarray(order, dtype=np.int64)


def to_circuit_order(grad_heis: np.ndarray, order: np.ndarray) -> np.ndarray:
    out = np.zeros(len(order), dtype=np.float64)
    out[order] = grad_heis
    return out


def run_adjoint(kind: int, n_units: int, observable: Sequence[Tuple[int, float]],
                layers: Sequence[Sequence[Tuple[int, float]]], max_weight: Optional[int] = None,
                min_coeff: Optional[float] = None, min_terms: Optional[int] = None,
                noise_table: Optional[Callable[[int, int], Sequence[float]]] = None, diag_state: int = 0,
                checkpoint_every: Optional[int] = 4, engine: Optional[DeviceEngine] = None,
                tape_budget: int = 64 << 30, sparse: Optional[bool] = None) -> GradientRun:
    """Expectation value and d<O>/dtheta for every rotation of ``layers``
    (circuit order, each a list of (generator key, angle)).

    sparse=False records every gate on the sorted chained kernels (their row
    order is fixed, so the gradient is the same bit for bit for any checkpoint
    interval or tape budget); by default gates record on the sparse passes,
    whose gradient sums follow the order their lists were filled in and agree
    to rounding (about 1e-16 relative)."""
    eng = engine if engine is not None else DeviceEngine(kind, n_units)
    if sparse is not None:
        eng.set_sparse(sparse)
    W = eng.W
    eng.set_cutoff(max_weight, min_coeff, min_terms)
    heis, order = heisenberg_layers(layers, W)
    step = None
    if noise_table is not None:
        def step(e, li, nl):
            e.scale_by_weight(noise_table(li, nl), reclaim=True)
    sess = AdjointSession(eng, keys_to_array([k for k, _ in observable], W),
                          np.array([c for _, c in observable], dtype=np.float64), heis, step, diag_state,
                          checkpoint_every, tape_budget)
    ev = sess.forward()
    g = sess.backward()
    return GradientRun(ev, to_circuit_order(g, order), sess.n_terms, sess.forward_s, sess.backward_s,
                       sess.term_gate_updates, sess.tape_bytes_peak, len(sess.segments) - 1)


_AUTOGRAD = None


def autograd_function():
    """``PropagatedExpectation``, a ``torch.autograd.Function`` (built on
    first use, so importing the package does not import torch): forward runs
    the device propagation with checkpoints, backward the reverse sweep."""
    global _AUTOGRAD
    if _AUTOGRAD is None:
        import torch

        class PropagatedExpectation(torch.autograd.Function):
            @staticmethod
            def forward(ctx, angles, make_session):
                sess = make_session(angles.detach().to("cpu", torch.float64).numpy())
                value = sess.forward()
                ctx.sess = sess
                return torch.tensor(value, dtype=torch.float64, device=angles.device)

            @staticmethod
            def backward(ctx, grad_out):
                sess = ctx.sess
                g = to_circuit_order(sess.backward(), sess.order)
                ctx.sess = None
                sess.eng.close()
                return grad_out * torch.from_numpy(g).to(grad_out.device), None

        _AUTOGRAD = PropagatedExpectation
    return _AUTOGRAD


def expectation(propagator, observable, circuit, angles, initial_state=None, checkpoint_every: Optional[int] = 4):
    """Differentiable <O>: ``angles`` (torch tensor, one entry per rotation
    of ``circuit`` in circuit order) replace the rotations' angles.  Returns a
    0-d float64 tensor whose backward gives d<O>/dangle."""
    fn = autograd_function()
    n_rot = sum(len(layer) for layer in circuit.layers)
    if angles.numel() != n_rot:
        raise ValueError(f"propaq_amd: {angles.numel()} angles given for {n_rot} rotations")

    def make_session(theta: np.ndarray):
        return propagator._adjoint_session(observable, circuit, theta, initial_state, checkpoint_every)

    return fn.apply(angles.reshape(-1), make_session)
