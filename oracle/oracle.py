"""ctypes front end of the FP64 CPU oracle (test infrastructure only).

The oracle restates the reference's single-partition propagation semantics in
plain C (``oracle/pq_oracle.c``).  Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg import this module, and only as the
checker; the product engine in ``propaq_amd`` never touches it.

``PartitionedOracle`` is the same engine split over S hash partitions run on
threads (``PartitionedTermSum``, ``crates/core/src/engine/partitioned_termsum.rs``),
the multi-core CPU baseline; its results equal the single partition's key for
key.

``run_oracle`` mirrors the run driver of ``crates/pauli/src/engine.rs:74-238``
(and its Majorana twin ``crates/majorana/src/engine.rs:74``): seed admission,
layers and gates in reverse (Heisenberg) order, per-layer noise then reclaim,
per-gate live counts, expectation with the deferred frame applied, and
``terms_below``.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np

PAULI, MAJORANA, TOY = 0, 1, 2

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_lib", "libpq_oracle.so")
_lib = None


def build() -> str:
    """Compile the C oracle with its Makefile (a few seconds)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = ctypes.CDLL(_LIB_PATH)
    P, U64P, DP = ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_double)
    sig = {
        "pqo_create": (P, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "pqo_part_create": (P, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "pqo_part_destroy": (None, [P]),
        "pqo_part_set_cutoff": (None, [P, ctypes.c_int, ctypes.c_uint32, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_int, ctypes.c_uint64]),
        "pqo_part_set_defer": (None, [P, ctypes.c_int]),
        "pqo_part_len": (ctypes.c_size_t, [P]),
        "pqo_part_admits_initial": (ctypes.c_int, [P, U64P, ctypes.c_double]),
        "pqo_part_add": (None, [P, U64P, ctypes.c_double]),
        "pqo_part_apply_rotations": (None, [P, U64P, DP, ctypes.c_size_t, U64P]),
        "pqo_part_scale_by_table": (None, [P, DP, ctypes.c_size_t]),
        "pqo_part_reclaim": (ctypes.c_size_t, [P]),
        "pqo_part_expectation": (ctypes.c_double, [P, U64P, ctypes.c_int]),
        "pqo_part_terms_below": (ctypes.c_size_t, [P, ctypes.c_double]),
        "pqo_part_export": (None, [P, U64P, DP]),
        "pqo_part_counters": (None, [P, U64P, DP]),
        "pqo_part_owner": (ctypes.c_int, [P, U64P]),
        "pqo_part_partition_len": (ctypes.c_size_t, [P, ctypes.c_int]),
        "pqo_destroy": (None, [P]),
        "pqo_set_cutoff": (None, [P, ctypes.c_int, ctypes.c_uint32, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_int, ctypes.c_uint64]),
        "pqo_set_defer": (None, [P, ctypes.c_int]),
        "pqo_set_truncator": (None, [P, P, P, ctypes.c_int]),
        "pqo_set_layer": (None, [P, ctypes.c_uint32, ctypes.c_uint32]),
        "pqo_len": (ctypes.c_size_t, [P]),
        "pqo_frame_is_identity": (ctypes.c_int, [P]),
        "pqo_add": (None, [P, U64P, ctypes.c_double]),
        "pqo_admits_initial": (ctypes.c_int, [P, U64P, ctypes.c_double]),
        "pqo_weight": (ctypes.c_uint32, [ctypes.c_int, U64P, ctypes.c_int, ctypes.c_int]),
        "pqo_trace": (ctypes.c_double, [ctypes.c_int, U64P, ctypes.c_int, ctypes.c_int, U64P, ctypes.c_int]),
        "pqo_product": (ctypes.c_int, [ctypes.c_int, U64P, U64P, ctypes.c_int, ctypes.c_int, U64P]),
        "pqo_anticommutes": (ctypes.c_int, [ctypes.c_int, U64P, U64P, ctypes.c_int, ctypes.c_int]),
        "pqo_apply_rotation": (ctypes.c_size_t, [P, U64P, ctypes.c_double]),
        "pqo_apply_rotations": (None, [P, U64P, DP, ctypes.c_size_t, U64P]),
        "pqo_scale_by_table": (None, [P, DP, ctypes.c_size_t]),
        "pqo_scale_by_factors": (None, [P, DP]),
        "pqo_reclaim": (ctypes.c_size_t, [P, ctypes.c_char_p]),
        "pqo_expectation": (ctypes.c_double, [P, U64P, ctypes.c_int]),
        "pqo_terms_below": (ctypes.c_size_t, [P, ctypes.c_double]),
        "pqo_export_raw": (None, [P, U64P, DP]),
        "pqo_export": (None, [P, U64P, DP]),
        "pqo_frame_conjugate": (ctypes.c_double, [P, U64P, U64P]),
        "pqo_frame_conjugate_generator": (ctypes.c_double, [P, U64P, U64P]),
        "pqo_frame_compose_rotation": (ctypes.c_int, [P, U64P, ctypes.c_double, ctypes.POINTER(ctypes.c_int)]),
        "pqo_counters": (None, [P, U64P, DP]),
        "pqo_tape_eval": (ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint8),
                                         ctypes.POINTER(ctypes.c_uint32), U64P, U64P, DP, ctypes.c_size_t, DP,
                                         ctypes.POINTER(ctypes.c_int64), ctypes.c_size_t, ctypes.c_size_t,
                                         ctypes.c_size_t, DP, DP]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def words_for(n_units: int) -> int:
    """Width ladder of ``crates/pauli/src/engine.rs:289-299``: 1, 2, 4, ... words."""
    w = 1
    while 32 * w < n_units:
        w *= 2
    return w


def to_words(key: int, W: int) -> np.ndarray:
    return np.array([(key >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(W)], dtype=np.uint64)


def from_words(words) -> int:
    v = 0
    for i, w in enumerate(words):
        v |= int(w) << (64 * i)
    return v


def _u64p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class OracleEngine:
    """One single-partition oracle term sum over ``n_units`` qubits or sites."""

    def __init__(self, kind: int, n_units: int, W: Optional[int] = None):
        self.kind, self.n_units = kind, n_units
        self.W = W if W is not None else words_for(n_units)
        self._h = lib().pqo_create(kind, n_units, self.W)
        if not self._h:
            raise ValueError("oracle: bad kind or width")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().pqo_destroy(h)
            self._h = None

    def set_cutoff(self, max_weight=None, min_coeff=None, min_terms=None):
        lib().pqo_set_cutoff(self._h, int(max_weight is not None), int(max_weight or 0),
                             int(min_coeff is not None), float(min_coeff or 0.0),
                             int(min_terms is not None), int(min_terms or 0))

    def set_defer(self, on: bool):
        lib().pqo_set_defer(self._h, int(bool(on)))

    def set_truncator(self, keep_fn, ctx, reads_key: bool):
        """A truncation plugin's scalar ``propaq_truncator_keep`` (a C function
        pointer, as an int or ctypes function) and its context pointer."""
        fp = ctypes.cast(keep_fn, ctypes.c_void_p).value if not isinstance(keep_fn, int) else keep_fn
        lib().pqo_set_truncator(self._h, fp, ctx, int(bool(reads_key)))

    def set_layer(self, layer_index: int, n_layers: int):
        lib().pqo_set_layer(self._h, int(layer_index), int(n_layers))

    def __len__(self):
        return int(lib().pqo_len(self._h))

    @property
    def frame_is_identity(self) -> bool:
        return bool(lib().pqo_frame_is_identity(self._h))

    def add(self, key: int, coeff: float):
        k = to_words(key, self.W)
        lib().pqo_add(self._h, _u64p(k), float(coeff))

    def admits_initial(self, key: int, coeff: float) -> bool:
        k = to_words(key, self.W)
        return bool(lib().pqo_admits_initial(self._h, _u64p(k), float(coeff)))

    def apply_rotation(self, gen: int, angle: float) -> int:
        g = to_words(gen, self.W)
        return int(lib().pqo_apply_rotation(self._h, _u64p(g), float(angle)))

    def apply_rotations(self, gens: np.ndarray, angles: np.ndarray) -> np.ndarray:
        """gens uint64[count, W], angles float64[count]; live counts after each."""
        g = np.ascontiguousarray(gens, dtype=np.uint64)
        a = np.ascontiguousarray(angles, dtype=np.float64)
        out = np.zeros(len(a), dtype=np.uint64)
        if len(a):
            lib().pqo_apply_rotations(self._h, _u64p(g), _dp(a), len(a), _u64p(out))
        return out

    def scale_by_table(self, table: Sequence[float]):
        t = np.ascontiguousarray(table, dtype=np.float64)
        lib().pqo_scale_by_table(self._h, _dp(t), t.size)

    def scale_by_factors(self, factors: np.ndarray):
        f = np.ascontiguousarray(factors, dtype=np.float64)
        assert f.size == len(self)
        lib().pqo_scale_by_factors(self._h, _dp(f))

    def reclaim(self, keep: Optional[np.ndarray] = None) -> int:
        if keep is None:
            return int(lib().pqo_reclaim(self._h, None))
        flags = np.ascontiguousarray(keep, dtype=np.uint8)
        assert flags.size == len(self)
        return int(lib().pqo_reclaim(self._h, flags.ctypes.data_as(ctypes.c_char_p)))

    def expectation(self, diag: int = 0) -> float:
        nd = max(1, (self.n_units + 63) // 64)
        d = to_words(diag, nd)
        return float(lib().pqo_expectation(self._h, _u64p(d), nd))

    def terms_below(self, thr: float) -> int:
        return int(lib().pqo_terms_below(self._h, float(thr)))

    def export(self, frame: bool = True):
        """(keys uint64[n, W], coeffs float64[n]) in row order."""
        n = len(self)
        keys = np.zeros((n, self.W), dtype=np.uint64)
        coeffs = np.zeros(n, dtype=np.float64)
        if n:
            (lib().pqo_export if frame else lib().pqo_export_raw)(self._h, _u64p(keys), _dp(coeffs))
        return keys, coeffs

    def terms(self, frame: bool = True) -> dict:
        keys, coeffs = self.export(frame)
        return {from_words(k): float(c) for k, c in zip(keys, coeffs)}

    def frame_conjugate(self, key: int, generator_direction: bool = False):
        k = to_words(key, self.W)
        out = np.zeros(self.W, dtype=np.uint64)
        f = lib().pqo_frame_conjugate_generator if generator_direction else lib().pqo_frame_conjugate
        s = f(self._h, _u64p(k), _u64p(out))
        return from_words(out), float(s)

    def frame_compose_rotation(self, gen: int, angle: float):
        g = to_words(gen, self.W)
        cw = ctypes.c_int(0)
        ok = lib().pqo_frame_compose_rotation(self._h, _u64p(g), float(angle), ctypes.byref(cw))
        return bool(ok), bool(cw.value)

    def counters(self) -> dict:
        c = np.zeros(4, dtype=np.uint64)
        d = np.zeros(2, dtype=np.float64)
        lib().pqo_counters(self._h, _u64p(c), _dp(d))
        return {"visited": int(c[0]), "emitted": int(c[1]), "declined": int(c[2]), "hits": int(c[3]),
                "discarded_coeff_l1": float(d[0]), "discarded_coeff_max": float(d[1])}


class PartitionedOracle:
    """The oracle split over ``partitions`` hash partitions, run on ``threads``
    threads (PartitionedTermSum, partitioned_termsum.rs:276-350): the
    multi-core CPU baseline.  Supports what the benchmark runs (cutoffs,
    Clifford deferral, weight tables and reclaim); truncation plugins and
    per-row factors use the single-partition OracleEngine."""

    def __init__(self, kind: int, n_units: int, W: Optional[int] = None, partitions: int = 1, threads: int = 1):
        self.kind, self.n_units = kind, n_units
        self.W = W if W is not None else words_for(n_units)
        self.partitions, self.threads = int(partitions), int(threads)
        self._h = lib().pqo_part_create(kind, n_units, self.W, self.partitions, self.threads)
        if not self._h:
            raise ValueError("oracle: bad kind, width or partition count")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().pqo_part_destroy(h)
            self._h = None

    def set_cutoff(self, max_weight=None, min_coeff=None, min_terms=None):
        lib().pqo_part_set_cutoff(self._h, int(max_weight is not None), int(max_weight or 0),
                                  int(min_coeff is not None), float(min_coeff or 0.0),
                                  int(min_terms is not None), int(min_terms or 0))

    def set_defer(self, on: bool):
        lib().pqo_part_set_defer(self._h, int(bool(on)))

    def __len__(self):
        return int(lib().pqo_part_len(self._h))

    def add(self, key: int, coeff: float):
        lib().pqo_part_add(self._h, _u64p(to_words(key, self.W)), float(coeff))

    def admits_initial(self, key: int, coeff: float) -> bool:
        return bool(lib().pqo_part_admits_initial(self._h, _u64p(to_words(key, self.W)), float(coeff)))

    def owner(self, key: int) -> int:
        return int(lib().pqo_part_owner(self._h, _u64p(to_words(key, self.W))))

    def partition_lens(self) -> list:
        return [int(lib().pqo_part_partition_len(self._h, q)) for q in range(self.partitions)]

    def apply_rotations(self, gens: np.ndarray, angles: np.ndarray) -> np.ndarray:
        g = np.ascontiguousarray(gens, dtype=np.uint64)
        a = np.ascontiguousarray(angles, dtype=np.float64)
        out = np.zeros(len(a), dtype=np.uint64)
        if len(a):
            lib().pqo_part_apply_rotations(self._h, _u64p(g), _dp(a), len(a), _u64p(out))
        return out

    def apply_rotation(self, gen: int, angle: float) -> int:
        self.apply_rotations(to_words(gen, self.W).reshape(1, self.W), np.array([angle]))
        return len(self)

    def scale_by_table(self, table: Sequence[float]):
        t = np.ascontiguousarray(table, dtype=np.float64)
        lib().pqo_part_scale_by_table(self._h, _dp(t), t.size)

    def reclaim(self, keep: Optional[np.ndarray] = None) -> int:
        if keep is not None:
            raise ValueError("partitioned oracle: reclaim by keep flags is not supported")
        return int(lib().pqo_part_reclaim(self._h))

    def expectation(self, diag: int = 0) -> float:
        nd = max(1, (self.n_units + 63) // 64)
        return float(lib().pqo_part_expectation(self._h, _u64p(to_words(diag, nd)), nd))

    def terms_below(self, thr: float) -> int:
        return int(lib().pqo_part_terms_below(self._h, float(thr)))

    def export(self, frame: bool = True):
        if not frame:
            raise ValueError("partitioned oracle: raw export is not supported")
        n = len(self)
        keys = np.zeros((n, self.W), dtype=np.uint64)
        coeffs = np.zeros(n, dtype=np.float64)
        if n:
            lib().pqo_part_export(self._h, _u64p(keys), _dp(coeffs))
        return keys, coeffs

    def terms(self, frame: bool = True) -> dict:
        keys, coeffs = self.export(frame)
        return {from_words(k): float(c) for k, c in zip(keys, coeffs)}

    def counters(self) -> dict:
        c = np.zeros(4, dtype=np.uint64)
        d = np.zeros(2, dtype=np.float64)
        lib().pqo_part_counters(self._h, _u64p(c), _dp(d))
        return {"visited": int(c[0]), "emitted": int(c[1]), "declined": int(c[2]), "hits": int(c[3]),
                "discarded_coeff_l1": float(d[0]), "discarded_coeff_max": float(d[1])}


def weight(kind: int, key: int, n_units: int, W: Optional[int] = None) -> int:
    W = W or words_for(n_units)
    k = to_words(key, W)
    return int(lib().pqo_weight(kind, _u64p(k), W, n_units))


def trace(kind: int, key: int, n_units: int, diag: int, W: Optional[int] = None) -> float:
    W = W or words_for(n_units)
    k = to_words(key, W)
    nd = max(1, (n_units + 63) // 64)
    d = to_words(diag, nd)
    return float(lib().pqo_trace(kind, _u64p(k), W, n_units, _u64p(d), nd))


def product(kind: int, gen: int, key: int, n_units: int, W: Optional[int] = None):
    """(G*P key, phase exponent e) with phase = i**e."""
    W = W or words_for(n_units)
    g, k, out = to_words(gen, W), to_words(key, W), np.zeros(W, dtype=np.uint64)
    e = int(lib().pqo_product(kind, _u64p(g), _u64p(k), W, n_units, _u64p(out)))
    return from_words(out), e


def anticommutes(kind: int, gen: int, key: int, n_units: int, W: Optional[int] = None) -> bool:
    W = W or words_for(n_units)
    g, k = to_words(gen, W), to_words(key, W)
    return bool(lib().pqo_anticommutes(kind, _u64p(g), _u64p(k), W, n_units))


@dataclass
class OracleResult:
    expectation_value: float
    n_terms: list = field(default_factory=list)
    terms_below_cutoff: int = 0
    terms: Optional[dict] = None
    counters: Optional[dict] = None
    per_gate: Optional[list] = None
    term_gate_updates: int = 0


def run_oracle(kind: int, n_units: int, observable: Sequence, layers: Sequence[Sequence],
               max_weight=None, min_coeff=None, min_terms=None,
               noise_table: Optional[Callable[[int, int], Sequence[float]]] = None,
               noise_factors: Optional[Callable] = None,
               diag_state: Optional[int] = 0, defer: bool = True, want_terms: bool = False,
               gate_hook: Optional[Callable] = None, W: Optional[int] = None,
               truncator: Optional[tuple] = None, partitions: Optional[int] = None,
               threads: int = 1) -> OracleResult:
    """Run the reference driver on the oracle.

    observable: [(key:int, coeff)]; layers: [[(gen:int, angle)]] in circuit
    order (reversed here, as ``extract_layers`` does); noise_table(layer_idx,
    n_layers) -> factor table indexed by weight, or None for no noise;
    noise_factors(engine, layer_idx, n_layers) -> per-row factors for a
    key-aware model.  truncator = (keep_fn, ctx, reads_key) of a truncation
    plugin (``EmitCutoff::term``); the driver then sets the layer context
    before each layer and turns Clifford deferral off for a key-reading plugin
    (``crates/pauli/src/engine.rs:101-103, 118``).  partitions (with threads):
    run the PartitionedOracle instead of the single-partition engine.
    """
    if partitions is not None:
        if truncator is not None or noise_factors is not None:
            raise ValueError("partitioned oracle: truncation plugins and per-row factors are not supported")
        eng = PartitionedOracle(kind, n_units, W, partitions, threads)
    else:
        eng = OracleEngine(kind, n_units, W)
    eng.set_cutoff(max_weight, min_coeff, min_terms)
    if truncator is not None:
        eng.set_truncator(*truncator)
        defer = defer and not truncator[2]
    eng.set_defer(defer)
    for key, c in observable:
        if eng.admits_initial(key, c):
            eng.add(key, c)
    heis = [list(reversed(layer)) for layer in reversed(layers)]
    n_layers = len(heis)
    n_terms = []
    updates = 0
    for li, layer in enumerate(heis):
        if truncator is not None:
            eng.set_layer(li, n_layers)
        if noise_table is not None or noise_factors is not None:
            if noise_factors is not None:
                eng.scale_by_factors(noise_factors(eng, li, n_layers))
            else:
                eng.scale_by_table(noise_table(li, n_layers))
            eng.reclaim()
        if gate_hook is None:  # the whole layer in one call
            gens = np.array([to_words(g, eng.W) for g, _ in layer], dtype=np.uint64).reshape(len(layer), eng.W)
            before = len(eng)
            after = eng.apply_rotations(gens, np.array([a for _, a in layer], dtype=np.float64))
            counts = [int(v) for v in after]
            updates += before + sum(counts[:-1]) if counts else 0
            n_terms.extend(counts)
            continue
        for gen, angle in layer:
            updates += len(eng)
            eng.apply_rotation(gen, angle)
            n_terms.append(len(eng))
            gate_hook(eng)
    ev = eng.expectation(diag_state) if diag_state is not None else 0.0
    below = eng.terms_below(min_coeff) if min_coeff is not None else 0
    return OracleResult(ev, n_terms, below, eng.terms() if want_terms else None, eng.counters(),
                        term_gate_updates=updates)


def tape_eval(kinds, params, a, b, vals, overlaps, roots, n_params: int, param_sets) -> np.ndarray:
    """The C restatement of a surrogate tape's evaluation, one set after
    another on one core (``pqo_tape_eval``); roots < 0 are empty."""
    k = np.ascontiguousarray(kinds, dtype=np.uint8)
    p = np.ascontiguousarray(params, dtype=np.uint32)
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    v = np.ascontiguousarray(vals, dtype=np.float64)
    ov = np.ascontiguousarray(overlaps, dtype=np.float64)
    r = np.ascontiguousarray(roots, dtype=np.int64)
    sets = np.ascontiguousarray(np.atleast_2d(np.asarray(param_sets, dtype=np.float64)))
    out = np.zeros(sets.shape[0], dtype=np.float64)
    rc = lib().pqo_tape_eval(k.size, k.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                             p.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), _u64p(a), _u64p(b), _dp(v), ov.size,
                             _dp(ov), r.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), int(n_params),
                             sets.shape[0], sets.shape[1], _dp(sets), _dp(out))
    if rc != 0:
        raise MemoryError("oracle: tape evaluation out of memory")
    return out
