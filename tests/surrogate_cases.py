"""Seeded surrogate workloads shared by the oracle pins and the device parity tests."""

from __future__ import annotations

import math
import random

from oracle import surrogate as S

CONFIGS = [
    S.Config(),
    S.Config(frequency=3),
    S.Config(coefficient=0.05),
    S.Config(weight=2),
    S.Config(simplify=True),
    S.Config(frequency=4, simplify=True, min_terms=6),
    S.Config(coefficient=0.01, frequency=5, weight=3),
]


def random_case(seed: int, kind: int, n: int = 4, n_gates: int = 12, n_params: int = 5, n_obs: int = 3,
                clifford: float = 0.15, numeric: float = 0.15):
    """(observable, layers) with symbolic, numeric and Clifford-angle gates."""
    rng = random.Random(seed)
    obs = [(rng.getrandbits(2 * n) or 1, rng.uniform(-1, 1)) for _ in range(n_obs)]
    layers = []
    for _ in range(n_gates):
        g = rng.getrandbits(2 * n) or 2
        r = rng.random()
        if r < clifford:
            layers.append([(g, None, rng.choice([math.pi / 2, -math.pi / 2, math.pi]))])
        elif r < clifford + numeric:
            layers.append([(g, None, rng.uniform(-3, 3))])
        else:
            layers.append([(g, rng.randrange(n_params), None)])
    return obs, layers


def random_params(seed: int, n_sets: int, n_params: int):
    rng = random.Random(seed)
    return [[rng.uniform(-math.pi, math.pi) for _ in range(n_params)] for _ in range(n_sets)]


def numeric_layers(layers, params):
    """The same circuit with every symbolic gate bound to params."""
    return [[(g, params[p] if p is not None else a) for g, p, a in layer] for layer in layers]


def wide_case(kind: int, n: int, seed: int, n_gates: int = 14):
    """Generators on five units spread over every key word (n_units 40, 100,
    200 give key widths 2, 4, 8), a numeric angle every fifth gate."""
    rng = random.Random(seed)
    active = [0, n // 3, n // 2, (2 * n) // 3, n - 1]

    def sparse(k):
        key = 0
        for _ in range(k):
            key |= rng.choice([1, 2, 3]) << (2 * rng.choice(active))
        return key

    obs = [(sparse(2) or 2, rng.uniform(-1, 1)) for _ in range(3)]
    layers = [[(sparse(2) or 1, i % 4 if i % 5 else None, None if i % 5 else rng.uniform(-3, 3))]
              for i in range(n_gates)]
    return obs, layers
