"""Batched propagation: many parameter sets (and observables) of one circuit
structure in one device store.

The reference runs parameter sets one after another; its only batching API,
``SurrogateModel::evaluate_batch`` (``crates/surrogate/src/model.rs:80-102``),
evaluates a precompiled symbolic tape, a different algorithm.  Here every
instance's rows live in one sorted store, tagged with the instance id in the
top bits of the key (``pq_engine_create_batched``), and each gate is one
launch sequence for all instances with per-instance angles.  Uneven depths are
aligned at the circuit's end: in Heisenberg order instance b applies its layer
L_b - 1 - t at step t, so every instance starts at once and an instance
finishes after its own L_b steps; its expectation value is read then and its
rows are removed.  The layers applied at one step must have the same
generators for every live instance (true of hardware-efficient ansatz
instances that differ in angles and depth).

Per instance the result is the run of that instance alone, bit for bit
(same keys, coefficients and summation order), except that a Clifford angle
is deferred into the frame only when every live instance shares it.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from ._native import DeviceEngine, words_to_int
from .runner import keys_to_array

Rot = Tuple[int, float]


@dataclass
class BatchRun:
    expectation_values: np.ndarray          # per instance
    n_terms: List[int] = field(default_factory=list)  # total live rows after each batched gate
    final_counts: Optional[np.ndarray] = None         # per instance, live rows when it finished
    term_gate_updates: int = 0
    seconds: float = 0.0
    steps: int = 0
    terms: Optional[List[dict]] = None       # per instance {key: coeff} (want_terms)


def schedule(instances: Sequence[Sequence[Sequence[Rot]]], W: int):
    """Heisenberg-order steps: (generators uint64[n, W], angles float64[n, B])
    per step, and the number of steps of every instance."""
    B = len(instances)
    depth = np.array([len(layers) for layers in instances], dtype=np.int64)
    steps = []
    for t in range(int(depth.max(initial=0))):
        live = [b for b in range(B) if depth[b] > t]
        ref = list(reversed(instances[live[0]][depth[live[0]] - 1 - t]))
        gens = [g for g, _ in ref]
        angles = np.zeros((len(gens), B), dtype=np.float64)
        for b in live:
            layer = list(reversed(instances[b][depth[b] - 1 - t]))
            if [g for g, _ in layer] != gens:
                raise ValueError(f"propaq_amd: instance {b} applies different generators than instance {live[0]} "
                                 f"at Heisenberg step {t}")
            angles[:, b] = [a for _, a in layer]
        steps.append((keys_to_array(gens, W), angles))
    return steps, depth


def run_batch(kind: int, n_units: int, observables, instances: Sequence[Sequence[Sequence[Rot]]],
              max_weight: Optional[int] = None, min_coeff: Optional[float] = None,
              noise_table: Optional[Sequence[float]] = None, diag_states=0, defer: bool = True,
              want_terms: bool = False, engine: Optional[DeviceEngine] = None) -> BatchRun:
    """Expectation value of every instance.  ``observables`` is one
    observable [(key, coeff)] for all instances or one per instance;
    ``instances`` holds each instance's circuit-order layers of (generator key,
    angle); ``diag_states`` one basis state or one per instance; the same
    ``noise_table`` (factor by weight) acts at every layer start."""
    B = len(instances)
    eng = engine if engine is not None else DeviceEngine(kind, n_units, n_instances=B)
    if eng.n_instances != B:
        raise ValueError("propaq_amd: engine instance count differs from the batch")
    W = eng.W
    eng.set_cutoff(max_weight, min_coeff, None)
    eng.set_defer_cliffords(defer)
    single = len(observables) == 0 or isinstance(observables[0][0], (int, np.integer))
    inst, keys, coeffs = [], [], []
    for b in range(B):
        for k, c in (observables if single else observables[b]):
            inst.append(b)
            keys.append(k)
            coeffs.append(c)
    eng.seed_batched(np.array(inst, dtype=np.uint32), keys_to_array(keys, W), np.array(coeffs, dtype=np.float64))
    diags = [int(diag_states)] * B if np.isscalar(diag_states) else [int(v) for v in diag_states]
    steps, depth = schedule(instances, W)
    values = np.zeros(B)
    final = np.zeros(B, dtype=np.uint64)
    terms = [None] * B if want_terms else None
    n_terms: List[int] = []
    c0 = eng.counters()["term_gate_updates"]
    t0 = time.perf_counter()

    def finish(which):
        flags = np.zeros(B, dtype=np.uint8)
        flags[list(which)] = 1
        if want_terms:
            split = split_terms(eng)
            for b in which:
                terms[b] = split.get(b, {})
        # the finished instances' values and counts, then their rows removed
        ev, counts = eng.readout_retire(diags, flags)
        for b in which:
            values[b] = ev[b]
            final[b] = counts[b]

    done0 = [b for b in range(B) if depth[b] == 0]
    if done0:
        finish(done0)
    for t, (gens, angles) in enumerate(steps):
        if noise_table is not None:
            eng.scale_by_weight(noise_table, reclaim=True)
        if len(gens):
            n_terms.extend(int(v) for v in eng.apply_gates_batched(gens, angles.reshape(len(gens), B)))
        done = [b for b in range(B) if depth[b] == t + 1]
        if done:
            finish(done)
    eng.synchronize()
    secs = time.perf_counter() - t0
    run = BatchRun(values, n_terms, final, eng.counters()["term_gate_updates"] - c0, secs, len(steps), terms)
    return run


@dataclass
class BatchGradientRun:
    expectation_values: np.ndarray     # per instance
    gradients: List[np.ndarray]        # per instance: d<O_b>/dtheta in its circuit order (layers, then gates)
    forward_s: float = 0.0
    backward_s: float = 0.0            # recomputation included
    term_gate_updates: int = 0         # of the forward run
    tape_bytes_peak: int = 0
    checkpoints: int = 0
    steps: int = 0


def run_batch_gradient(kind: int, n_units: int, observables, instances: Sequence[Sequence[Sequence[Rot]]],
                       max_weight: Optional[int] = None, min_coeff: Optional[float] = None, diag_states=0,
                       checkpoint_every: Optional[int] = 4, tape_budget: int = 64 << 30,
                       engine: Optional[DeviceEngine] = None, sparse: Optional[bool] = None) -> BatchGradientRun:
    """Expectation value and d<O>/dtheta of every rotation, for every
    instance, in one batched forward run and one reverse sweep.  The schedule
    is that of ``run_batch`` except that finished instances are frozen
    instead of being removed (removal is not reversible), so all values are
    read at the end.

    sparse=False records every gate on the sorted chained kernels: each
    instance's gradient then equals that of ``adjoint.run_adjoint(...,
    sparse=False)`` on the instance alone, bit for bit (its rows are reversed
    in the blocks of that run).  By default the gates record on the sparse
    passes (much faster; up to 4096 instances), whose list order varies from
    run to run: values are the same, gradients agree to rounding."""
    from .adjoint import BatchAdjointSession

    B = len(instances)
    eng = engine if engine is not None else DeviceEngine(kind, n_units, n_instances=B)
    if eng.n_instances != B:
        raise ValueError("propaq_amd: engine instance count differs from the batch")
    if sparse is not None:
        eng.set_sparse(sparse)
    W = eng.W
    eng.set_cutoff(max_weight, min_coeff, None)
    single = len(observables) == 0 or isinstance(observables[0][0], (int, np.integer))
    inst, keys, coeffs = [], [], []
    for b in range(B):
        for k, c in (observables if single else observables[b]):
            inst.append(b)
            keys.append(k)
            coeffs.append(c)
    diags = [int(diag_states)] * B if np.isscalar(diag_states) else [int(v) for v in diag_states]
    steps, depth = schedule(instances, W)
    sess = BatchAdjointSession(eng, np.array(inst, dtype=np.uint32), keys_to_array(keys, W),
                               np.array(coeffs, dtype=np.float64), steps, diags, depth, checkpoint_every, tape_budget)
    values = sess.forward()
    g = sess.backward()  # [Heisenberg gates, B]
    offs = np.cumsum([0] + [len(a) for _, a in steps])
    grads = []
    for b in range(B):
        layers = instances[b]
        starts = np.cumsum([0] + [len(l) for l in layers])
        out = np.zeros(int(starts[-1]), dtype=np.float64)
        for t in range(int(depth[b])):
            li = int(depth[b]) - 1 - t
            n_g = len(layers[li])
            # step t applies layer li last gate first
            out[starts[li]:starts[li] + n_g] = g[offs[t]:offs[t] + n_g, b][::-1]
        grads.append(out)
    return BatchGradientRun(values, grads, sess.forward_s, sess.backward_s, sess.term_gate_updates,
                            sess.tape_bytes_peak, max(len(sess.segments) - 1, 0), len(steps))


def split_terms(eng: DeviceEngine) -> dict:
    """{instance: {key: coeff}} of the live rows, frame applied."""
    keys, coeffs = eng.export(frame=True)
    shift, W = eng.instance_shift, eng.W
    mask = (1 << shift) - 1 if shift < 64 else (1 << 64) - 1
    out: dict = {}
    for k, c in zip(keys, coeffs):
        b = int(k[W - 1]) >> shift
        words = list(k)
        words[W - 1] = int(words[W - 1]) & mask
        out.setdefault(b, {})[words_to_int(words)] = float(c)
    return out
