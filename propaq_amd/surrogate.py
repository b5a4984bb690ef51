"""Symbolic surrogate propagation: build once, evaluate for many parameter sets.

Mirrors the reference's surrogate API (``crates/surrogate/src/propagator.rs``,
``model.rs``, ``truncation.rs`` and ``propaq/circuits/*/surrogate_*.py``):
``PauliSurrogatePropagator(truncation=...).build(observable, circuit,
initial_state)`` returns a ``PauliSurrogateModel`` whose ``evaluate(params)``
and ``evaluate_batch(param_sets)`` give the expectation value at any angles.

The build runs in the native library on the host (``pq_surrogate_*``); the
compiled model lives in HBM and is evaluated by the gfx950 tape kernels for a
whole batch of parameter sets at once (``pq_model_evaluate``).  There is no
other path: without the library or a device these calls raise.
"""

from __future__ import annotations

import gzip
import struct
from typing import List, Optional, Sequence

import numpy as np

from . import _native
from .circuits import MajoranaCircuit, PauliCircuit
from .events import write_surrogate_log
from .plugins import NativeTruncator
from .runner import keys_to_array
from .truncation import (CoefficientTruncator, FrequencyTruncator, Simplify, TermBudget, Truncator,
                         WeightTruncator, resolve_truncation)

MAX_UNITS = 256


# --------------------------------------------------------------- circuits

class SurrogateRotation:
    """A rotation with either a symbolic parameter index or a numeric angle
    (``propaq/circuits/pauli/surrogate_rotation.py``): exactly one is given."""

    def __init__(self, generator, param_index: Optional[int] = None, angle: Optional[float] = None,
                 is_intermediate: bool = False, qiskit_gate_idx: Optional[int] = None):
        if (param_index is None) == (angle is None):
            raise ValueError(
                "SurrogateRotation requires exactly one of `param_index` (symbolic) or `angle` (numeric, baked in "
                f"at build time), got param_index={param_index!r}, angle={angle!r}")
        if param_index is not None and int(param_index) < 0:
            raise ValueError("param_index must be non-negative")
        self.generator = generator
        self.param_index = None if param_index is None else int(param_index)
        self.angle = None if angle is None else float(angle)
        self.is_intermediate = is_intermediate
        self.qiskit_gate_idx = qiskit_gate_idx

    def __repr__(self):
        what = f"param_index={self.param_index}" if self.param_index is not None else f"angle={self.angle}"
        return f"{type(self).__name__}({self.generator!r}, {what})"


class SurrogateMajoranaRotation(SurrogateRotation):
    pass


class _SurrogateCircuit:
    _rotation = SurrogateRotation

    def __init__(self, layers: List[list]):
        self._layers = layers
        self.parameter_sources: list = []
        self.qiskit_parameters: tuple = ()

    @property
    def layers(self) -> List[list]:
        return self._layers

    @property
    def rotations(self) -> list:
        return [r for layer in self._layers for r in layer]

    @property
    def n_params(self) -> int:
        """Number of distinct parameter indices (max param_index + 1)."""
        idx = [r.param_index for r in self.rotations if r.param_index is not None]
        return max(idx) + 1 if idx else 0

    @classmethod
    def _from_circuit(cls, circuit, param_indices: Sequence[Optional[int]], **kw):
        rotations = circuit.rotations
        if len(param_indices) != len(rotations):
            raise ValueError(f"param_indices length ({len(param_indices)}) must match circuit.rotations length "
                             f"({len(rotations)})")
        flat, layers = 0, []
        for layer in circuit.layers:
            new = []
            for rot in layer:
                idx = param_indices[flat]
                new.append(cls._rotation(rot.generator, param_index=idx, angle=None if idx is not None else rot.angle,
                                         is_intermediate=getattr(rot, "is_intermediate", False),
                                         qiskit_gate_idx=getattr(rot, "qiskit_gate_idx", None)))
                flat += 1
            layers.append(new)
        return cls(layers, **kw)

    @classmethod
    def from_generators_and_param_indices(cls, generators, param_indices, **kw):
        if len(generators) != len(param_indices):
            raise ValueError("generators and param_indices must have equal length")
        return cls([[cls._rotation(g, i)] for g, i in zip(generators, param_indices)], **kw)


class SurrogatePauliCircuit(_SurrogateCircuit):
    """``propaq/circuits/pauli/surrogate_circuit.py``."""

    @classmethod
    def from_pauli_circuit(cls, circuit: PauliCircuit, param_indices: Sequence[Optional[int]]):
        return cls._from_circuit(circuit, param_indices)


class SurrogateMajoranaCircuit(_SurrogateCircuit):
    """``propaq/circuits/majorana/surrogate_circuit.py``."""

    _rotation = SurrogateMajoranaRotation

    def __init__(self, layers: List[list], n_modes: Optional[int] = None):
        super().__init__(layers)
        self.n_modes = n_modes

    @classmethod
    def from_majorana_circuit(cls, circuit: MajoranaCircuit, param_indices: Sequence[Optional[int]]):
        return cls._from_circuit(circuit, param_indices, n_modes=getattr(circuit, "n_modes", None))

    @classmethod
    def from_generators_and_param_indices(cls, generators, param_indices, n_modes: Optional[int] = None):
        return super().from_generators_and_param_indices(generators, param_indices, n_modes=n_modes)


# ------------------------------------------------------------- truncation

class FrequencyTruncationPolicy:
    """Legacy surrogate policy (``crates/surrogate/src/truncation.rs``)."""

    def __init__(self, max_frequency: Optional[int] = None, weight_cutoff: Optional[int] = None,
                 min_terms: Optional[int] = None):
        self.max_frequency = max_frequency
        self.weight_cutoff = weight_cutoff
        self.min_terms = min_terms

    def decompose(self) -> List[Truncator]:
        ops: List[Truncator] = []
        if self.max_frequency is not None:
            ops.append(FrequencyTruncator(self.max_frequency))
        if self.weight_cutoff is not None:
            ops.append(WeightTruncator(self.weight_cutoff))
        if self.min_terms is not None:
            ops.append(TermBudget(self.min_terms))
        return ops

    def __repr__(self):
        return (f"FrequencyTruncationPolicy(max_frequency={self.max_frequency}, weight_cutoff={self.weight_cutoff}, "
                f"min_terms={self.min_terms})")


def _resolve(truncation) -> List[Truncator]:
    """propagator.rs:24-31 + truncators.rs:178-185: the legacy policy, the
    shared forms, and no NativeTruncator."""
    if isinstance(truncation, FrequencyTruncationPolicy):
        ops = truncation.decompose()
    elif isinstance(truncation, (FrequencyTruncator, Simplify)):
        ops = [truncation]
    elif isinstance(truncation, (list, tuple)) and all(isinstance(t, Truncator) for t in truncation):
        ops = list(truncation)
    else:
        ops = resolve_truncation(truncation)
    if any(isinstance(t, NativeTruncator) for t in ops):
        raise ValueError("NativeTruncator only applies to numerical propagation, use it with "
                         "PauliPropagator/MajoranaPropagator instead")
    return ops


def _config(ops: List[Truncator]) -> dict:
    """ResolvedConfig (truncators.rs:187-211): last one of each kind wins."""
    c = {"frequency": None, "coefficient": None, "weight": None, "min_terms": None, "simplify": False}
    for t in ops:
        if isinstance(t, FrequencyTruncator):
            c["frequency"] = None if t.frequency is None else int(t.frequency)
        elif isinstance(t, CoefficientTruncator):
            c["coefficient"] = t.coefficient
        elif isinstance(t, WeightTruncator):
            c["weight"] = t.weight
        elif isinstance(t, TermBudget):
            c["min_terms"] = t.min_terms
        elif isinstance(t, Simplify):
            c["simplify"] = bool(t.enabled)
    return c


# ----------------------------------------------------------------- models

_MAGIC = struct.unpack("<I", b"PQSM")[0]  # model.rs:268
_FORMAT_VERSION = 9
_OP_SIZE = {0: 8, 1: 16, 2: 16, 3: 12, 4: 12}


class _SurrogateModel:
    """A compiled surrogate in device memory (``model.rs:44-262``)."""

    def __init__(self, device_model: "_native.DeviceModel"):
        self._m = device_model

    @property
    def n_params(self) -> int:
        return self._m.n_params

    @property
    def n_terms(self) -> int:
        return self._m.n_terms

    @property
    def n_monomials(self) -> int:
        return self._m.n_monomials

    def _rows(self, param_sets) -> np.ndarray:
        rows = [list(map(float, p)) for p in param_sets]
        for i, p in enumerate(rows):
            if len(p) < self.n_params:
                raise ValueError(f"param_sets[{i}] has {len(p)} elements but model requires {self.n_params}")
        width = max([self.n_params] + [len(p) for p in rows]) if rows else self.n_params
        out = np.zeros((len(rows), max(width, 1)), dtype=np.float64)
        for i, p in enumerate(rows):
            out[i, :len(p)] = p
        return out

    def evaluate(self, params: Sequence[float]) -> float:
        if len(params) < self.n_params:
            raise ValueError(f"params has {len(params)} elements but model requires {self.n_params}")
        return float(self._m.evaluate(self._rows([params]))[0])

    def evaluate_batch(self, param_sets) -> List[float]:
        if isinstance(param_sets, np.ndarray) and param_sets.ndim == 2:
            if param_sets.shape[1] < self.n_params:
                raise ValueError(f"param_sets rows have {param_sets.shape[1]} elements but model requires "
                                 f"{self.n_params}")
            if param_sets.shape[0] == 0:
                return []
            return [float(v) for v in self._m.evaluate(param_sets)]
        if len(param_sets) == 0:
            return []
        return [float(v) for v in self._m.evaluate(self._rows(param_sets))]

    # --- file format of model.rs:124-172 (PQSM, version 9): gzip blocks of the
    # serialized tape (symcoeff.rs:950-980) and of (overlap f64, root u64) terms

    def save(self, path: str) -> None:
        kinds, params, a, b, vals, overlaps, roots = self._m.export()
        n = kinds.size
        size = 8 + sum(1 + _OP_SIZE[int(k)] for k in kinds)
        raw = bytearray(size)
        struct.pack_into("<Q", raw, 0, n)
        pos = 8
        for i in range(n):
            k = int(kinds[i])
            raw[pos] = k
            pos += 1
            if k == 0:
                struct.pack_into("<d", raw, pos, float(vals[i]))
            elif k == 1:
                struct.pack_into("<QQ", raw, pos, int(a[i]), int(b[i]))
            elif k == 2:
                struct.pack_into("<dQ", raw, pos, float(vals[i]), int(a[i]))
            else:
                struct.pack_into("<IQ", raw, pos, int(params[i]), int(a[i]))
            pos += _OP_SIZE[k]
        terms = np.zeros(overlaps.size, dtype=[("ov", "<f8"), ("root", "<u8")])
        terms["ov"], terms["root"] = overlaps, roots
        tape_blob = gzip.compress(bytes(raw))
        term_blob = gzip.compress(terms.tobytes())
        with open(path, "wb") as fh:
            fh.write(struct.pack("<IIQ", _MAGIC, _FORMAT_VERSION, self.n_params))
            fh.write(struct.pack("<QQ", 1, len(tape_blob)))
            fh.write(tape_blob)
            fh.write(struct.pack("<QQ", 1, len(term_blob)))
            fh.write(term_blob)

    @classmethod
    def load(cls, path: str):
        with open(path, "rb") as fh:
            data = fh.read()
        try:
            magic, version, n_params = struct.unpack_from("<IIQ", data, 0)
        except struct.error as e:
            raise OSError("unrecognized or outdated surrogate model file (format changed); rebuild the model") from e
        if magic != _MAGIC or version != _FORMAT_VERSION:
            raise OSError("unrecognized or outdated surrogate model file (format changed); rebuild the model")
        pos = 16

        def blobs():
            nonlocal pos
            (count,) = struct.unpack_from("<Q", data, pos)
            pos += 8
            lens = struct.unpack_from(f"<{count}Q", data, pos)
            pos += 8 * count
            out = []
            for ln in lens:
                out.append(gzip.decompress(data[pos:pos + ln]))
                pos += ln
            return out

        kinds, params, a, b, vals = [], [], [], [], []
        for raw in blobs():
            (n,) = struct.unpack_from("<Q", raw, 0)
            p = 8
            for _ in range(n):
                k = raw[p]
                p += 1
                if k == 0:
                    (v,) = struct.unpack_from("<d", raw, p)
                    kinds.append(0), params.append(0), a.append(0), b.append(0), vals.append(v)
                elif k == 1:
                    x, y = struct.unpack_from("<QQ", raw, p)
                    kinds.append(1), params.append(0), a.append(x), b.append(y), vals.append(0.0)
                elif k == 2:
                    v, x = struct.unpack_from("<dQ", raw, p)
                    kinds.append(2), params.append(0), a.append(x), b.append(0), vals.append(v)
                elif k in (3, 4):
                    q, x = struct.unpack_from("<IQ", raw, p)
                    kinds.append(k), params.append(q), a.append(x), b.append(0), vals.append(0.0)
                else:
                    raise OSError(f"corrupt surrogate tape: unknown op tag {k}")
                p += _OP_SIZE[k]
        term_raw = b"".join(blobs())
        terms = np.frombuffer(term_raw, dtype=[("ov", "<f8"), ("root", "<u8")])
        dm = _native.DeviceModel.from_tape(int(n_params), kinds, params, a, b, vals, terms["ov"].copy(),
                                           terms["root"].copy())
        return cls(dm)

    def __repr__(self):
        return f"{type(self).__name__}(n_terms={self.n_terms}, n_params={self.n_params})"


class PauliSurrogateModel(_SurrogateModel):
    pass


class MajoranaSurrogateModel(_SurrogateModel):
    pass


# ------------------------------------------------------------ variational

class ParamSource:
    """One surrogate parameter slot (``propaq/circuits/_qiskit_symbolic.py:13``):
    ``scale * value_of(parameter)``, or the constant ``scale`` when
    ``parameter`` is None."""

    def __init__(self, parameter, scale: float = 1.0):
        self.parameter = parameter
        self.scale = float(scale)

    def __repr__(self):
        return f"ParamSource(parameter={self.parameter!r}, scale={self.scale})"


class VariationalSurrogateModel:
    """A compiled surrogate addressed by named parameters instead of raw slots
    (the role of ``propaq/models/variational.py``).

    The slot list is compiled once into a gather plan: slot k reads
    ``parameters[pos[k]]`` scaled by ``scale[k]``, or is the constant
    ``scale[k]`` when ``pos[k] < 0``.  An assignment (a mapping keyed by the
    parameter objects, or values in the order of ``parameters``) becomes one
    raw vector through that plan; a batch becomes one matrix and a single
    device evaluation.  Parameters may be any hashable objects."""

    def __init__(self, model: "_SurrogateModel", parameter_sources: Sequence[ParamSource], parameters: tuple):
        self._model = model
        self.parameters = tuple(parameters)
        where = {p: i for i, p in enumerate(self.parameters)}
        srcs = list(parameter_sources)
        self._pos = np.array([-1 if s.parameter is None else where[s.parameter] for s in srcs], dtype=np.int64)
        self._scale = np.array([s.scale for s in srcs], dtype=np.float64)
        self._const = self._pos < 0

    @property
    def n_params(self) -> int:
        """Raw parameter slots of the compiled model (a parameter may feed several)."""
        return int(self._pos.size)

    def _named_values(self, values) -> np.ndarray:
        from collections.abc import Mapping

        if isinstance(values, Mapping):
            return np.array([values[p] for p in self.parameters], dtype=np.float64)
        v = np.asarray(values, dtype=np.float64).reshape(-1)
        if v.size != len(self.parameters):
            raise ValueError(f"expected {len(self.parameters)} parameter values (one per distinct parameter), "
                             f"got {v.size}")
        return v

    def raw_vector(self, values) -> List[float]:
        """The raw slot vector the compiled model is evaluated at."""
        return self._raw(self._named_values(values)).tolist()

    def _raw(self, named: np.ndarray) -> np.ndarray:
        picked = named[np.where(self._const, 0, self._pos)] if named.size else np.zeros(self._pos.size)
        return np.where(self._const, self._scale, self._scale * picked)

    def evaluate(self, values) -> float:
        return self._model.evaluate(self.raw_vector(values))

    def evaluate_batch(self, values_batch) -> List[float]:
        raw = np.array([self._raw(self._named_values(v)) for v in values_batch]).reshape(-1, self.n_params)
        # the device model takes the matrix as is; any other model gets rows
        return self._model.evaluate_batch(raw if isinstance(self._model, _SurrogateModel) else raw.tolist())


# ------------------------------------------------------------ propagators

class _SurrogatePropagator:
    _kind = 0
    _model_cls = _SurrogateModel

    def __init__(self, truncation=None, n_threads: Optional[int] = None, logger=None, progress_bar: bool = False,
                 progress_every: int = 1):
        self._truncators = _resolve(truncation)
        self._logger = logger
        self.n_threads = n_threads
        self.progress_bar = progress_bar
        self.progress_every = max(1, int(progress_every))
        self.last_n_terms: List[int] = []

    @property
    def truncators(self) -> List[Truncator]:
        return list(self._truncators)

    def set_truncation(self, truncation=None) -> None:
        self._truncators = _resolve(truncation)

    def _units(self, observable) -> int:
        raise NotImplementedError

    def build(self, observable, circuit, initial_state=None):
        """propagator.rs:163-222: propagate symbolically, truncate after every
        gate, compile against the basis state initial_state (default 0)."""
        n_units = self._units(observable)
        if n_units > MAX_UNITS:
            raise ValueError(f"propaq_amd: {n_units} units exceeds the surrogate's maximum of {MAX_UNITS}")
        n_params = int(circuit.n_params)
        b = _native.SurrogateBuilder(self._kind, max(n_units, 1))
        try:
            b.set_truncation(**_config(self._truncators))
            b.seed(observable.key_array(b.W), observable.coeff_array())
            layers = list(reversed(circuit.layers))
            gates = [r for layer in layers for r in reversed(layer)]
            gens = keys_to_array([r.generator.key for r in gates], b.W)
            pidx = np.array([-1 if r.param_index is None else int(r.param_index) for r in gates], dtype=np.int64)
            angles = np.array([0.0 if r.angle is None else float(r.angle) for r in gates], dtype=np.float64)
            if self._logger is None:
                self.last_n_terms = [int(v) for v in b.apply(gens, pidx, angles)] if gates else []
            else:
                stats = b.apply_logged(gens, pidx, angles) if gates else []
                self.last_n_terms = [st["terms"] for st in stats]
                layer_of = [li for li, layer in enumerate(layers) for _ in layer]
                recs = [dict(st, gate_idx=i, layer_idx=layer_of[i],
                             qiskit_gate_idx=getattr(gates[i], "qiskit_gate_idx", None)) for i, st in enumerate(stats)]
                ctr = b.counters()
                share = lambda a, c: a / c if c else 0.0  # noqa: E731
                phases = {"partitions": 1, "scan_s": ctr["rotation_s"], "absorb_s": 0.0, "claims_s": 0.0,
                          "scan_occupancy": None, "absorb_occupancy": None, "claims_occupancy": None,
                          "terms": len(b), "inline_positions": 64 * b.W, "overflow_rows": 0, "overflow_share": 0.0,
                          "visited": ctr["visited"], "emitted": ctr["emitted"], "declined": 0,
                          "emitted_share": share(ctr["emitted"], ctr["visited"]), "declined_share": 0.0,
                          "exchange_hits": ctr["hits"], "exchange_hit_share": share(ctr["hits"], ctr["emitted"])}
                write_surrogate_log(self._logger, recs, _config(self._truncators), len(observable), phases)
            state = 0 if initial_state is None else int(initial_state)
            model = b.compile(state, n_params)
        finally:
            b.close()
        return self._model_cls(model)


class PauliSurrogatePropagator(_SurrogatePropagator):
    """Back-propagates Pauli observables symbolically (``propagator.rs:126-243``)."""

    _kind = _native.BASIS_PAULI
    _model_cls = PauliSurrogateModel

    def _units(self, observable) -> int:
        return observable.n_units


class MajoranaSurrogatePropagator(_SurrogatePropagator):
    """Back-propagates Majorana observables symbolically (``propagator.rs:245-360``)."""

    _kind = _native.BASIS_MAJORANA
    _model_cls = MajoranaSurrogateModel

    def _units(self, observable) -> int:
        n_modes = observable.n_units
        if n_modes % 2:
            raise ValueError("propaq_amd: n_modes must be even")
        return n_modes // 2
