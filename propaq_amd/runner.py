"""Run driver over the device engine.

Mirrors ``run_at_width`` (``crates/pauli/src/engine.rs:74-238``, Majorana twin
``crates/majorana/src/engine.rs:74``): the observable is admitted and seeded,
layers are applied last-first with their gates last-first (Heisenberg order,
``extract_layers`` :44-70), every layer starts with the noise pass and a
reclaim, and the live count is recorded after every gate.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

from ._native import DeviceEngine, int_to_words, words_to_int


def keys_to_array(keys: Sequence[int], W: int) -> np.ndarray:
    out = np.zeros((len(keys), W), dtype=np.uint64)
    for i, k in enumerate(keys):
        out[i] = int_to_words(int(k), W)
    return out


def gates_to_arrays(rots: Sequence[Tuple[int, float]], W: int):
    gens = keys_to_array([g for g, _ in rots], W)
    angles = np.array([float(a) for _, a in rots], dtype=np.float64)
    return gens, angles


@dataclass
class DeviceRun:
    expectation_value: float
    n_terms: List[int] = field(default_factory=list)
    terms_below_cutoff: int = 0
    keys: Optional[np.ndarray] = None
    coeffs: Optional[np.ndarray] = None
    counters: Optional[dict] = None
    sparse_key_bytes: int = 0

    def terms(self) -> dict:
        assert self.keys is not None
        return {words_to_int(k): float(c) for k, c in zip(self.keys, self.coeffs)}


def run_device(kind: int, n_units: int, observable: Sequence[Tuple[int, float]], layers: Sequence[Sequence],
               max_weight: Optional[int] = None, min_coeff: Optional[float] = None,
               min_terms: Optional[int] = None,
               noise_table: Optional[Callable[[int, int], Sequence[float]]] = None,
               diag_state: Optional[int] = 0, defer: bool = True, want_terms: bool = False,
               layer_hook: Optional[Callable] = None, engine: Optional[DeviceEngine] = None,
               truncator=None, small_runs: Optional[bool] = None, sparse: Optional[bool] = None,
               pairs: Optional[bool] = None) -> DeviceRun:
    """``truncator``: a ``NativeTruncator`` deciding branches, seeds and
    reclaims (``EmitCutoff::term``); Clifford deferral is then off when it
    reads keys (``engine.rs:101-103``)."""
    eng = engine if engine is not None else DeviceEngine(kind, n_units)
    W = eng.W
    eng.set_cutoff(max_weight, min_coeff, min_terms)
    if small_runs is not None:
        eng.set_small_runs(small_runs)
    if sparse is not None:
        eng.set_sparse(sparse)
    if pairs is not None:
        eng.set_pairs(pairs)
    if truncator is not None:
        eng.set_keep_kernel(truncator.keep_flags)
        defer = defer and not truncator.reads_key
    eng.set_defer_cliffords(defer)
    eng.seed(keys_to_array([k for k, _ in observable], W), np.array([c for _, c in observable], dtype=np.float64))
    heis = [list(reversed(layer)) for layer in reversed(layers)]
    n_layers = len(heis)
    n_terms: List[int] = []
    if noise_table is None and layer_hook is None and truncator is None:
        # nothing happens between layers: one call applies them all
        heis = [[g for layer in heis for g in layer]]
    for li, layer in enumerate(heis):
        if truncator is not None:
            eng.set_layer(li, n_layers)
        if noise_table is not None:
            eng.scale_by_weight(noise_table(li, n_layers), reclaim=True)
        if layer_hook is not None:
            layer_hook(eng, li, n_layers)
        gens, angles = gates_to_arrays(layer, W)
        n_terms.extend(int(v) for v in eng.apply_gates(gens, angles))
    ev = eng.expectation(diag_state) if diag_state is not None else 0.0
    below = eng.terms_below(min_coeff) if min_coeff is not None else 0
    run = DeviceRun(ev, n_terms, below, counters=eng.counters(), sparse_key_bytes=eng.key_bytes())
    if want_terms:
        run.keys, run.coeffs = eng.export(frame=True)
    return run


def uniform_table(damping: float, n_units: int) -> List[float]:
    """UniformNoiseModel's table exp(-damping w), w in [0, n_units]
    (crates/core/src/engine/noise_resolver.rs:87-94)."""
    return [math.exp(-damping * w) for w in range(n_units + 1)]
