"""Synthetic circuits and observables of the five benchmark workloads.

Keys use the interleaved layout of ``crates/core/src/algebra/strings.rs:30-41``
(unit u owns bits 2u, 2u+1; Pauli bit 2q = X, 2q+1 = Z; Majorana bit k =
gamma_k).  Rotations are (generator key, angle) with the engine's
exp(-i angle G / 2) convention; circuits are lists of layers in circuit order
(the driver reverses them, as ``crates/pauli/src/engine.rs:44-70`` does).

Gate decompositions follow the reference's helpers:
  * CZ = CP(pi): Z_i (pi/2), Z_j (pi/2), Z_i Z_j (-pi/2)
    (``propaq/datatypes/pauli/termsum.py:34-44``);
  * Majorana XX+YY(theta, beta=0) on neighbours i<j:
    gamma_2i gamma_2j+1 at +theta/2 and gamma_2i+1 gamma_2j at -theta/2
    (``propaq/datatypes/majorana/termsum.py:19-44``);
  * Majorana CP(phi): n_i and n_j monomials at -phi/2, the 4-mode monomial
    at +phi/2 (``propaq/datatypes/majorana/termsum.py:57-67``);
  * Majorana observable n_p = (I - Z_p)/2 becomes 0.5 I + 0.5 gamma_2p gamma_2p+1
    through the reference's JW transform (``majorana/termsum.py:164-197,390-415``).

Conventions fixed here (they change results, so oracle and device share them):
  * 2D lattices are open, site (r, c) -> 10 r + c, edges in row-major order
    (right neighbour, then down neighbour);
  * one layer per gate family per step (noise is applied at every layer start);
  * the Majorana density-density block is scheduled as 31 rounds of a
    round-robin tournament over the 32 sites, one layer per round;
  * CZ ladders are two brick layers (even bonds, then odd bonds);
  * the half-filled Majorana state occupies sites 0..15.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

PAULI, MAJORANA = 0, 1
Rot = Tuple[int, float]


def x(q: int) -> int:
    return 1 << (2 * q)


def z(q: int) -> int:
    return 1 << (2 * q + 1)


def y(q: int) -> int:
    return x(q) | z(q)


def pauli_key(xmask: int, zmask: int) -> int:
    """Interleave PauliString(x, z) bitmasks into an engine key
    (crates/pauli/src/algebra.rs:145-159)."""
    k, q = 0, 0
    while (xmask >> q) or (zmask >> q):
        if (xmask >> q) & 1:
            k |= x(q)
        if (zmask >> q) & 1:
            k |= z(q)
        q += 1
    return k


@dataclass
class Workload:
    name: str
    kind: int
    n_units: int
    observable: List[Tuple[int, float]]
    layers: List[List[Rot]]
    max_weight: Optional[int] = None
    min_coeff: Optional[float] = None
    damping: Optional[float] = None
    diag_state: int = 0
    params: dict = field(default_factory=dict)

    @property
    def n_gates(self) -> int:
        return sum(len(l) for l in self.layers)

    def noise_table(self):
        if self.damping is None:
            return None
        return [math.exp(-self.damping * w) for w in range(self.n_units + 1)]

    def truncated(self, steps_layers: int) -> "Workload":
        """The same workload restricted to its first ``steps_layers`` layers."""
        w = Workload(**{**self.__dict__})
        w.layers = self.layers[:steps_layers]
        w.name = f"{self.name}[:{steps_layers} layers]"
        return w


def kicked_ising_1d(n: int = 20, steps: int = 10, theta_zz: float = 0.3, theta_x: float = 0.4) -> List[List[Rot]]:
    layers = []
    for _ in range(steps):
        layers.append([(z(i) | z(i + 1), theta_zz) for i in range(n - 1)])
        layers.append([(x(i), theta_x) for i in range(n)])
    return layers


def lattice_edges(L: int = 10):
    edges = []
    for r in range(L):
        for c in range(L):
            s = r * L + c
            if c + 1 < L:
                edges.append((s, s + 1))
            if r + 1 < L:
                edges.append((s, s + L))
    return edges


def kicked_ising_2d(L: int = 10, steps: int = 20, theta_zz: float = -math.pi / 2, theta_x: float = 0.6,
                    seed: Optional[int] = None) -> List[List[Rot]]:
    """Kicked Ising on an open L x L lattice; with ``seed`` every angle is an
    independent draw from U[-pi, pi] (the differentiation workload)."""
    edges = lattice_edges(L)
    rng = np.random.default_rng(seed) if seed is not None else None
    layers = []
    for _ in range(steps):
        if rng is None:
            layers.append([(z(i) | z(j), theta_zz) for i, j in edges])
            layers.append([(x(i), theta_x) for i in range(L * L)])
        else:
            layers.append([(z(i) | z(j), float(rng.uniform(-math.pi, math.pi))) for i, j in edges])
            layers.append([(x(i), float(rng.uniform(-math.pi, math.pi))) for i in range(L * L)])
    return layers


def majorana_givens(theta: float, i: int, j: int) -> List[Rot]:
    lo, hi = min(i, j), max(i, j)
    d = hi - lo
    jw = 0
    for k in range(lo + 1, hi):
        jw |= (1 << (2 * k)) | (1 << (2 * k + 1))
    sign = 1 if ((d - 1) // 2) % 2 == 0 else -1
    f = theta / 2.0
    if i > j:
        m1 = (1 << (2 * hi)) | jw | (1 << (2 * lo + 1))
        m2 = (1 << (2 * hi + 1)) | jw | (1 << (2 * lo))
        return [(m1, -sign * f), (m2, sign * f)]
    m1 = (1 << (2 * lo)) | jw | (1 << (2 * hi + 1))
    m2 = (1 << (2 * lo + 1)) | jw | (1 << (2 * hi))
    return [(m1, sign * f), (m2, -sign * f)]


def majorana_cp(phi: float, i: int, j: int) -> List[Rot]:
    ni = (1 << (2 * i)) | (1 << (2 * i + 1))
    nj = (1 << (2 * j)) | (1 << (2 * j + 1))
    return [(ni, -phi / 2), (nj, -phi / 2), (ni | nj, phi / 2)]


def round_robin(n: int):
    """n - 1 rounds of n / 2 disjoint pairs covering every pair once."""
    ids = list(range(n))
    rounds = []
    for _ in range(n - 1):
        rounds.append([(min(ids[k], ids[n - 1 - k]), max(ids[k], ids[n - 1 - k])) for k in range(n // 2)])
        ids = [ids[0]] + [ids[-1]] + ids[1:-1]
    return rounds


def ucj_majorana(n_sites: int = 32, reps: int = 4, seed: int = 3) -> List[List[Rot]]:
    rng = np.random.default_rng(seed)
    layers = []
    for _ in range(reps):
        for layer in range(n_sites):
            start = layer % 2
            rots = []
            for i in range(start, n_sites - 1, 2):
                rots += majorana_givens(float(rng.uniform(-math.pi, math.pi)), i, i + 1)
            layers.append(rots)
        for rnd in round_robin(n_sites):
            rots = []
            for i, j in rnd:
                rots += majorana_cp(float(rng.normal(0.0, 0.5)), i, j)
            layers.append(rots)
    return layers


def hea(n: int, depth: int, rng) -> List[List[Rot]]:
    layers = []
    for _ in range(depth):
        layers.append([(y(i), float(rng.uniform(-math.pi, math.pi))) for i in range(n)])
        layers.append([(z(i), float(rng.uniform(-math.pi, math.pi))) for i in range(n)])
        for start in (0, 1):
            rots = []
            for i in range(start, n - 1, 2):
                rots += [(z(i), math.pi / 2), (z(i + 1), math.pi / 2), (z(i) | z(i + 1), -math.pi / 2)]
            layers.append(rots)
    return layers


def tfim_observable(n: int) -> List[Tuple[int, float]]:
    return [(z(i) | z(i + 1), 1.0) for i in range(n - 1)] + [(x(i), 0.5) for i in range(n)]


def config(idx: int) -> Workload:
    """The five benchmark workloads, 1-based as listed in the benchmark spec."""
    if idx == 1:
        return Workload("kicked-ising-1d-20q-10s", PAULI, 20, [(z(10), 1.0)], kicked_ising_1d(),
                        max_weight=8, min_coeff=1e-4)
    if idx == 2:
        return Workload("kicked-ising-2d-10x10-20s", PAULI, 100, [(z(55), 1.0)], kicked_ising_2d(),
                        max_weight=12, min_coeff=1e-5, damping=1e-3)
    if idx == 3:
        obs = [(0, 0.5 * 16)] + [((1 << (2 * p)) | (1 << (2 * p + 1)), 0.5) for p in range(16)]
        return Workload("majorana-ucj-32s-4r", MAJORANA, 32, obs, ucj_majorana(), max_weight=6,
                        min_coeff=1e-5, damping=1e-3, diag_state=(1 << 16) - 1)
    if idx == 4:
        return Workload("kicked-ising-2d-10x10-12s-random", PAULI, 100, [(z(55), 1.0)],
                        kicked_ising_2d(steps=12, seed=4), max_weight=10, min_coeff=1e-4)
    if idx == 5:
        rng = np.random.default_rng(5)
        depth = int(rng.integers(4, 17))
        return Workload("hea-50q-tfim", PAULI, 50, tfim_observable(50), hea(50, depth, rng), max_weight=10,
                        min_coeff=1e-5, params={"depth": depth})
    raise ValueError(f"no workload {idx}")


def hea_batch(n_instances: int = 256, seed: int = 5, n: int = 50, depth_lo: int = 4, depth_hi: int = 16):
    """Per-instance (depth, layers) of the batched HEA workload: depth from
    U{depth_lo..depth_hi}, then that many HEA layers, all from one seeded
    generator in instance order."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_instances):
        depth = int(rng.integers(depth_lo, depth_hi + 1))
        out.append((depth, hea(n, depth, rng)))
    return out
