"""Truncation API of the drop-in (``propaq/truncation``, Rust in
``crates/core/src/policy/truncators.rs``).

``resolve_truncation`` accepts what the reference's constructor accepts (a
list of truncators, one truncator, a legacy ``TruncationPolicy`` or None,
``truncators.rs:293-311``) and ``resolve_config`` collapses the list with
"last one wins" per kind (``truncators.rs:200-211``).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional


class Truncator:
    """Base of every truncator."""


class CoefficientTruncator(Truncator):
    """Coefficient cutoff: a branch whose |value| falls below ``coefficient`` is
    not formed (it is held back for the pair rule), and reclaim removes rows
    below it."""

    def __init__(self, coefficient: Optional[float] = None):
        self.coefficient = None if coefficient is None else float(coefficient)

    def __repr__(self):
        return f"CoefficientTruncator(coefficient={self.coefficient})"


class WeightTruncator(Truncator):
    """Drop whole terms whose weight exceeds ``weight``."""

    def __init__(self, weight: Optional[int] = None):
        self.weight = None if weight is None else int(weight)

    def __repr__(self):
        return f"WeightTruncator(weight={self.weight})"


class TermBudget(Truncator):
    """Live-term floor: below ``min_terms`` terms the lossy truncators are off."""

    def __init__(self, min_terms: Optional[int] = None):
        self.min_terms = None if min_terms is None else int(min_terms)

    def __repr__(self):
        return f"TermBudget(min_terms={self.min_terms})"


class FrequencyTruncator(Truncator):
    """Surrogate-only; rejected by the numerical propagators."""

    def __init__(self, frequency: Optional[int] = None):
        self.frequency = frequency


class Simplify(Truncator):
    """Surrogate-only; rejected by the numerical propagators."""

    def __init__(self, enabled: bool = True):
        self.enabled = enabled


class TruncationPolicy:
    """Legacy single-policy form, decomposed into truncators
    (truncators.rs:265-285)."""

    def __init__(self, weight_cutoff: Optional[int] = None, coeff_cutoff: float = 0.0,
                 min_terms: Optional[int] = None):
        self.weight_cutoff = weight_cutoff
        self.coeff_cutoff = float(coeff_cutoff)
        self.min_terms = min_terms

    def should_truncate(self, weight: int, abs_coeff: float) -> bool:
        return (self.weight_cutoff is not None and weight > self.weight_cutoff) or abs_coeff < self.coeff_cutoff

    def decompose(self) -> List[Truncator]:
        ops: List[Truncator] = []
        if self.weight_cutoff is not None:
            ops.append(WeightTruncator(self.weight_cutoff))
        if self.coeff_cutoff > 0.0:
            ops.append(CoefficientTruncator(self.coeff_cutoff))
        if self.min_terms is not None:
            ops.append(TermBudget(self.min_terms))
        return ops


@dataclass
class ResolvedConfig:
    weight: Optional[int] = None
    coefficient: Optional[float] = None
    min_terms: Optional[int] = None
    native: Optional[object] = None


def resolve_truncation(truncation) -> List[Truncator]:
    if truncation is None:
        return []
    if isinstance(truncation, TruncationPolicy):
        return truncation.decompose()
    if isinstance(truncation, Truncator):
        return [truncation]
    if isinstance(truncation, (list, tuple)) and all(isinstance(t, Truncator) for t in truncation):
        return list(truncation)
    raise TypeError("truncation must be a truncator (CoefficientTruncator/WeightTruncator/TermBudget/"
                    "NativeTruncator), a list of truncators, a TruncationPolicy, or None")


def reject_surrogate_only(truncators: List[Truncator]) -> None:
    if any(isinstance(t, (FrequencyTruncator, Simplify)) for t in truncators):
        raise ValueError("FrequencyTruncator/Simplify only apply to surrogate propagation, use "
                         "WeightTruncator / CoefficientTruncator / TermBudget with the numerical propagator")


def resolve_config(truncators: List[Truncator]) -> ResolvedConfig:
    from .plugins import NativeTruncator

    r = ResolvedConfig()
    for t in truncators:
        if isinstance(t, CoefficientTruncator):
            r.coefficient = t.coefficient
        elif isinstance(t, WeightTruncator):
            r.weight = t.weight
        elif isinstance(t, TermBudget):
            r.min_terms = t.min_terms
        elif isinstance(t, NativeTruncator):
            r.native = t
    return r
