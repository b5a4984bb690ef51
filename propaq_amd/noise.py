"""Noise models of the drop-in (``propaq/noise``, Rust in
``crates/core/src/policy/noise.rs:26-63`` and
``crates/core/src/engine/noise_resolver.rs``)."""

from __future__ import annotations

import math
from abc import ABC

from .plugins import NativeNoiseModel  # noqa: F401  (re-export)

PYTHON_TERM_HOOK = "damping_factor_term"  # noise_resolver.rs:23


class NoiseModel(ABC):
    """Base class of Python-side noise models.  A subclass supplies
    ``damping_factor(weight, active)`` (a per-weight table) or
    ``damping_factor_term(kind, words, n_units, weight)`` (per term)."""


class UniformNoiseModel(NoiseModel):
    r"""Each term of weight w is scaled by exp(-damping * w) at every layer."""

    def __init__(self, damping: float):
        self.damping = float(damping)

    def damping_factor(self, term_weight: int, active_modes: int = 0) -> float:
        return math.exp(-self.damping * term_weight)

    def table(self, max_weight: int):
        """noise_resolver.rs:87-94: exp(-damping w), w in [0, max_weight]."""
        return [math.exp(-self.damping * w) for w in range(max_weight + 1)]

    def apply_noise(self, term_sum) -> None:
        term_sum.apply_damping(self, 0)

    def __repr__(self):
        return f"UniformNoiseModel(damping={self.damping})"


class GateNoiseModel(NoiseModel):
    """A custom Python noise model: subclass it and define
    ``damping_factor(weight, active_modes)`` (tabulated once by weight) or
    ``damping_factor_term(basis_kind, words, n_units, weight)`` (called per
    term, key-aware; disables Clifford deferral)."""

    def __init__(self, *args, **kwargs):
        pass

    def damping_factor(self, _term_weight: int, _active_modes: int) -> float:
        return 1.0


NoiseModel.register(NativeNoiseModel)
