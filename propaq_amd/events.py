"""Structured JSON Lines event log of a propagation run, and its parser.

``Logger(filename, log_every=1)`` is the reference's ``Logger``
(``propaq/_rust_core/__init__.pyi:102-127``): given to a propagator, the run
writes one ``gate`` and one ``truncation`` record for every ``log_every``-th
gate and one closing ``engine_phases`` record, with the keys and meanings of
``write_gate_log`` / ``write_phase_event`` (``crates/pauli/src/propagator.rs:23-101``),
so the reference's ``LogParser`` reads these files unchanged.  The per-gate
figures come from the device counters (``pq_counters_get``): the propagator
applies a logged run one gate at a time, reads the counters after each gate
and resets them, so ``discarded_coeff_max`` is the gate's own maximum as in
``termsum.rs:611-626``.

``engine_phases`` keeps the reference's keys with their device meaning: one
partition (the GPU), ``scan_s`` = branch kernels, ``absorb_s`` = class
merges + store merge, ``claims_s`` = 0 (pair-rule claims happen inside the
store merge), occupancies null (there is no CPU pool), ``inline_positions`` =
the key bits a row holds (keys are never spilled, so ``overflow_rows`` = 0),
``exchange_hits`` = emitted branches that landed on a live key.

``LogParser`` reads such a file (or one written by the reference) into typed
events, with the accessors of the reference's ``propaq/log_parser.py``.
"""

from __future__ import annotations

import json
from dataclasses import dataclass, fields
from typing import List, Optional


class Logger:
    """JSON Lines event log for a propagator run; ``filename`` is overwritten."""

    def __init__(self, filename: str, log_every: int = 1):
        if int(log_every) < 1:
            raise ValueError("log_every must be >= 1")
        self._filename = str(filename)
        self._log_every = int(log_every)

    @property
    def filename(self) -> str:
        return self._filename

    @property
    def log_every(self) -> int:
        return self._log_every

    def __repr__(self):
        return f"Logger(filename={self._filename!r}, log_every={self._log_every})"


@dataclass
class GateRecord:
    """One applied gate (``crates/pauli/src/engine.rs:252-265``)."""

    gate_idx: int
    layer_idx: int
    qiskit_gate_idx: Optional[int]
    terms_before: int
    terms_after: int
    terms_gained: int
    declined: int
    discarded_coeff_l1: float
    discarded_coeff_max: float
    elapsed_ms: float


def _num(v) -> str:
    return "null" if v is None else str(v)


def _e(v: float, digits: int) -> str:
    """Rust's {:.Ne} float format (mantissa, 'e', exponent without padding)."""
    s = f"{float(v):.{digits}e}"
    mant, exp = s.split("e")
    return f"{mant}e{int(exp)}"


def write_log(logger: Logger, records: List[GateRecord], weight_cutoff: Optional[int],
              coeff_cutoff: Optional[float], phases: dict) -> None:
    """Write the run's records (every ``log_every``-th gate) and the closing
    ``engine_phases`` record."""
    every = max(1, logger.log_every)
    wc = _num(weight_cutoff)
    cc = _e(coeff_cutoff or 0.0, 6)
    with open(logger.filename, "w") as fh:
        for g in records:
            if g.gate_idx % every:
                continue
            q = _num(g.qiskit_gate_idx)
            fh.write(f'{{"event":"gate","gate_idx":{g.gate_idx},"layer_idx":{g.layer_idx},"qiskit_gate_idx":{q},'
                     f'"terms":{g.terms_after},"ms_per_gate":{_e(g.elapsed_ms, 3)}}}\n')
            fh.write(f'{{"event":"truncation","gate_idx":{g.gate_idx},"layer_idx":{g.layer_idx},'
                     f'"qiskit_gate_idx":{q},"trigger":"emit","truncation_model":"emit_gate",'
                     f'"terms_before":{g.terms_before},"terms_after":{g.terms_after},'
                     f'"terms_gained":{g.terms_gained},"terms_discarded":{g.declined},'
                     f'"discarded_coeff_l1":{_e(g.discarded_coeff_l1, 6)},'
                     f'"discarded_coeff_max":{_e(g.discarded_coeff_max, 6)},"weight_cutoff":{wc},'
                     f'"coeff_cutoff":{cc},"elapsed_ms":{_e(g.elapsed_ms, 3)}}}\n')
        fh.write(json.dumps(dict(event="engine_phases", **phases), separators=(",", ":")) + "\n")


def phases_record(terms: int, key_words: int, phase_ms: dict, totals: dict) -> dict:
    """The closing record from the profiled phase times and the run's counter totals."""
    visited, emitted, declined, hits = (int(totals.get(k, 0)) for k in ("visited", "emitted", "declined", "hits"))
    share = lambda a, b: a / b if b else 0.0  # noqa: E731
    return {
        "partitions": 1,
        "scan_s": phase_ms.get("branch", 0.0) / 1e3,
        "absorb_s": (phase_ms.get("child_sort", 0.0) + phase_ms.get("store_merge", 0.0)) / 1e3,
        "claims_s": 0.0,
        "scan_occupancy": None, "absorb_occupancy": None, "claims_occupancy": None,
        "terms": int(terms), "inline_positions": 64 * int(key_words), "overflow_rows": 0, "overflow_share": 0.0,
        "visited": visited, "emitted": emitted, "declined": declined,
        "emitted_share": share(emitted, visited), "declined_share": share(declined, visited),
        "exchange_hits": hits, "exchange_hit_share": share(hits, emitted),
    }


# ----------------------------------------------------------------- parser

@dataclass
class GateEvent:
    gate_idx: int
    layer_idx: int
    terms: int
    ms_per_gate: float
    qiskit_gate_idx: Optional[int] = None
    monomials: Optional[int] = None


@dataclass
class TruncationEvent:
    gate_idx: int
    layer_idx: int
    trigger: str
    terms_before: int
    terms_after: int
    terms_gained: int
    terms_discarded: int
    discarded_coeff_l1: float
    discarded_coeff_max: float
    weight_cutoff: Optional[int]
    coeff_cutoff: float
    elapsed_ms: float
    qiskit_gate_idx: Optional[int] = None


@dataclass
class SurrogateMergeEvent:
    """The truncation pass after one gate of a surrogate build
    (crates/surrogate/src/propagator.rs:123-157, propaq/log_parser.py:73)."""
    gate_idx: int
    layer_idx: int
    trigger: str
    terms_before: int
    terms_after: int
    terms_discarded: int
    monomials_before: int
    monomials_after: int
    monomials_discarded: int
    frequency: Optional[int]
    weight: Optional[int]
    coefficient: Optional[float]
    elapsed_ms: float
    qiskit_gate_idx: Optional[int] = None


@dataclass
class EnginePhasesEvent:
    partitions: int
    scan_s: float
    absorb_s: float
    claims_s: float
    scan_occupancy: Optional[float]
    absorb_occupancy: Optional[float]
    claims_occupancy: Optional[float]
    terms: int
    inline_positions: int
    overflow_rows: int
    overflow_share: float
    visited: int
    emitted: int
    declined: int
    emitted_share: float
    declined_share: float
    exchange_hits: int
    exchange_hit_share: float


def _build(cls, ev: dict):
    names = {f.name for f in fields(cls)}
    return cls(**{k: v for k, v in ev.items() if k in names})


class LogParser:
    """Typed view of a JSON Lines log (``propaq/log_parser.py:157``)."""

    def __init__(self, filename: str) -> None:
        self._filename = filename
        self.reload()

    def reload(self) -> None:
        self._gates: List[GateEvent] = []
        self._truncs: List[TruncationEvent] = []
        self._phases: List[EnginePhasesEvent] = []
        self._smerges: List[SurrogateMergeEvent] = []
        with open(self._filename) as fh:
            for line in fh:
                line = line.strip()
                if not line:
                    continue
                ev = json.loads(line)
                kind = ev.get("event")
                if kind == "gate":
                    self._gates.append(_build(GateEvent, ev))
                elif kind == "truncation":
                    self._truncs.append(_build(TruncationEvent, ev))
                elif kind == "engine_phases":
                    self._phases.append(_build(EnginePhasesEvent, ev))
                elif kind == "surrogate_merge":
                    self._smerges.append(_build(SurrogateMergeEvent, ev))

    @property
    def gate_events(self) -> List[GateEvent]:
        return list(self._gates)

    @property
    def truncation_events(self) -> List[TruncationEvent]:
        return list(self._truncs)

    @property
    def engine_phases_events(self) -> List[EnginePhasesEvent]:
        return list(self._phases)

    @property
    def gate_indices(self) -> List[int]:
        return [e.gate_idx for e in self._gates]

    @property
    def terms(self) -> List[int]:
        return [e.terms for e in self._gates]

    @property
    def ms_per_gate(self) -> List[float]:
        return [e.ms_per_gate for e in self._gates]

    @property
    def qiskit_gate_indices(self) -> List[Optional[int]]:
        return [e.qiskit_gate_idx for e in self._gates]

    @property
    def terms_before(self) -> List[int]:
        return [e.terms_before for e in self._truncs]

    @property
    def terms_after(self) -> List[int]:
        return [e.terms_after for e in self._truncs]

    @property
    def terms_gained(self) -> List[int]:
        return [e.terms_gained for e in self._truncs]

    @property
    def terms_discarded(self) -> List[int]:
        return [e.terms_discarded for e in self._truncs]

    @property
    def discarded_coeff_l1(self) -> List[float]:
        return [e.discarded_coeff_l1 for e in self._truncs]

    @property
    def discarded_coeff_max(self) -> List[float]:
        return [e.discarded_coeff_max for e in self._truncs]

    @property
    def elapsed_ms(self) -> List[float]:
        return [e.elapsed_ms for e in self._truncs]

    @property
    def surrogate_merge_events(self) -> List[SurrogateMergeEvent]:
        return list(self._smerges)

    @property
    def monomials_before(self) -> List[int]:
        return [e.monomials_before for e in self._smerges]

    @property
    def monomials_after(self) -> List[int]:
        return [e.monomials_after for e in self._smerges]

    @property
    def monomials_discarded(self) -> List[int]:
        return [e.monomials_discarded for e in self._smerges]


def write_surrogate_log(logger: Logger, gates: List[dict], cfg: dict, monomials_start: int, phases: dict) -> None:
    """A surrogate build's log (crates/surrogate/src/propagator.rs:94-160): every
    log_every-th gate's ``gate`` record, then its ``surrogate_merge`` records,
    then ``engine_phases``.  gates: per gate in execution order, gate_idx,
    layer_idx, qiskit_gate_idx and the builder's per-gate stats.
    monomials_before is the count after the previous gate's pass, as there."""
    every = max(1, logger.log_every)
    freq, weight = _num(cfg.get("frequency")), _num(cfg.get("weight"))
    coef = "null" if cfg.get("coefficient") is None else _e(cfg["coefficient"], 6)
    with open(logger.filename, "w") as fh:
        for g in gates:
            if g["gate_idx"] % every:
                continue
            fh.write(f'{{"event":"gate","gate_idx":{g["gate_idx"]},"layer_idx":{g["layer_idx"]},'
                     f'"qiskit_gate_idx":{_num(g["qiskit_gate_idx"])},"terms":{g["terms"]},'
                     f'"monomials":{g["monomials"]},"ms_per_gate":{_e(g["rotation_ms"], 3)}}}\n')
        before = monomials_start
        for g in gates:
            mono_before, before = before, g["monomials"]
            if g["gate_idx"] % every:
                continue
            fh.write(f'{{"event":"surrogate_merge","gate_idx":{g["gate_idx"]},"layer_idx":{g["layer_idx"]},'
                     f'"qiskit_gate_idx":{_num(g["qiskit_gate_idx"])},"trigger":"emit",'
                     f'"terms_before":{g["terms_rotated"]},"terms_after":{g["terms"]},'
                     f'"terms_discarded":{max(0, g["terms_rotated"] - g["terms"])},'
                     f'"monomials_before":{mono_before},"monomials_after":{g["monomials"]},'
                     f'"monomials_discarded":{max(0, mono_before - g["monomials"])},"frequency":{freq},'
                     f'"weight":{weight},"coefficient":{coef},"elapsed_ms":{_e(g["truncation_ms"], 3)}}}\n')
        fh.write(json.dumps(dict(event="engine_phases", **phases), separators=(",", ":")) + "\n")
