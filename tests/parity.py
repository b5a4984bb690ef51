"""Shared helpers for device-vs-oracle parity tests (imported by tests only)."""

from __future__ import annotations

import math
import os
import subprocess

import numpy as np

from oracle import oracle as orc
from propaq_amd import runner


def random_pauli_gen(rng, n: int, max_units: int = 3) -> int:
    k = 0
    units = rng.choice(n, size=int(rng.integers(1, max_units + 1)), replace=False)
    for u in units:
        p = int(rng.integers(1, 4))  # 1 = X, 2 = Z, 3 = Y
        if p & 1:
            k |= 1 << (2 * int(u))
        if p & 2:
            k |= 1 << (2 * int(u) + 1)
    return k


def random_majorana_gen(rng, n_sites: int, lengths=(2, 2, 4, 1, 3)) -> int:
    L = int(rng.choice(lengths))
    modes = rng.choice(2 * n_sites, size=L, replace=False)
    k = 0
    for m in modes:
        k |= 1 << int(m)
    return k


def random_angle(rng, clifford_frac=0.15, phase_frac=0.05) -> float:
    u = rng.random()
    if u < clifford_frac:
        return float(rng.choice([math.pi / 2, -math.pi / 2, 1.5 * math.pi]))
    if u < clifford_frac + phase_frac:
        return float(rng.choice([math.pi, -math.pi, 0.0]))
    return float(rng.uniform(-math.pi, math.pi))


def random_case(seed: int, kind: int, n_units: int, n_layers: int = 6, gates_per_layer: int = 8,
                n_obs: int = 3, clifford_frac=0.15, phase_frac=0.05):
    rng = np.random.default_rng(seed)
    gen = (lambda: random_pauli_gen(rng, n_units)) if kind == 0 else (lambda: random_majorana_gen(rng, n_units))
    layers = [[(gen(), random_angle(rng, clifford_frac, phase_frac)) for _ in range(gates_per_layer)]
              for _ in range(n_layers)]
    obs = [(gen(), float(rng.uniform(0.2, 1.5))) for _ in range(n_obs)]
    diag = int(rng.integers(0, 1 << min(n_units, 62)))
    return obs, layers, diag


def build_truncator(tmp_path, name, **defs) -> str:
    """Compile tests/truncator_plugin.c with gcc (-D options pick its rule)."""
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "truncator_plugin.c")
    so = os.path.join(str(tmp_path), f"{name}.so")
    flags = [f"-D{k}={v}" for k, v in defs.items()]
    subprocess.run(["gcc", "-shared", "-fPIC", "-O1", *flags, "-o", so, src, "-lm"], check=True)
    return so


def oracle_truncator(nt):
    """(scalar keep pointer, context, reads_key) of a loaded NativeTruncator."""
    import ctypes
    return ctypes.cast(nt._keep, ctypes.c_void_p).value, nt._ctx, nt.reads_key


_ORACLE_CACHE = {}


def oracle_run(cache_key, kind, n_units, obs, layers, *, max_weight=None, min_coeff=None, min_terms=None,
               damping=None, diag=0, defer=True, truncator=None):
    """The oracle's run of a case; with ``cache_key`` the result is kept for
    the other parametrisations of the same case (device paths differ, the
    oracle's answer does not)."""
    if cache_key is not None and cache_key in _ORACLE_CACHE:
        return _ORACLE_CACHE[cache_key]
    table = (lambda li, nl: runner.uniform_table(damping, n_units)) if damping is not None else None
    o = orc.run_oracle(kind, n_units, obs, layers, max_weight=max_weight, min_coeff=min_coeff,
                       min_terms=min_terms, noise_table=table, diag_state=diag, defer=defer, want_terms=True,
                       truncator=oracle_truncator(truncator) if truncator is not None else None)
    if cache_key is not None:
        _ORACLE_CACHE.clear()  # one case at a time: large term dicts are not kept around
        _ORACLE_CACHE[cache_key] = o
    return o


def compare(kind, n_units, obs, layers, *, max_weight=None, min_coeff=None, min_terms=None, damping=None,
            diag=0, defer=True, exact=True, ev_tol=1e-10, truncator=None, small_runs=None, cache_key=None,
            engine=None, sparse=None, pairs=None):
    table = (lambda li, nl: runner.uniform_table(damping, n_units)) if damping is not None else None
    o = oracle_run(cache_key, kind, n_units, obs, layers, max_weight=max_weight, min_coeff=min_coeff,
                   min_terms=min_terms, damping=damping, diag=diag, defer=defer, truncator=truncator)
    d = runner.run_device(kind, n_units, obs, layers, max_weight=max_weight, min_coeff=min_coeff,
                          min_terms=min_terms, noise_table=table, diag_state=diag, defer=defer,
                          want_terms=True, truncator=truncator, small_runs=small_runs, engine=engine,
                          sparse=sparse, pairs=pairs)
    assert d.n_terms == o.n_terms, "per-gate live counts differ"
    dt = d.terms()
    assert set(dt) == set(o.terms), "term keys differ"
    bad = [(k, dt[k], o.terms[k]) for k in dt if dt[k] != o.terms[k]]
    if exact:
        assert not bad, f"{len(bad)} coefficients differ, e.g. {bad[:3]}"
    else:
        for k, a, b in bad:
            assert abs(a - b) <= 1e-12 * max(abs(a), abs(b)), (k, a, b)
    assert abs(d.expectation_value - o.expectation_value) <= ev_tol, (d.expectation_value, o.expectation_value)
    assert d.terms_below_cutoff == o.terms_below_cutoff
    return d, o
