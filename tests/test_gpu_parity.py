"""Device engine vs the FP64 oracle: term keys, signs and coefficients
bit-exact after every run, per-gate live counts identical, expectation within
1e-10 absolute."""

import math

import pytest

from propaq_amd import workloads as wl
from tests.parity import compare, random_case

pytestmark = pytest.mark.gpu


def test_kicked_ising_1d_full_run_bit_exact():
    w = wl.config(1)
    compare(w.kind, w.n_units, w.observable, w.layers, max_weight=w.max_weight, min_coeff=w.min_coeff,
            diag=w.diag_state)


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("n_units", [6, 30])
def test_random_pauli_lossless(seed, n_units):
    obs, layers, diag = random_case(seed, 0, n_units, n_layers=5, gates_per_layer=6)
    compare(0, n_units, obs, layers, diag=diag)


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("n_units", [10, 40, 100])
def test_random_pauli_truncated_noisy(seed, n_units):
    obs, layers, diag = random_case(100 + seed, 0, n_units, n_layers=8, gates_per_layer=10)
    compare(0, n_units, obs, layers, max_weight=4, min_coeff=2e-3, damping=0.05, diag=diag)


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("n_units", [8, 24, 60])
def test_random_majorana_truncated_noisy(seed, n_units):
    obs, layers, diag = random_case(200 + seed, 1, n_units, n_layers=8, gates_per_layer=10)
    compare(1, n_units, obs, layers, max_weight=5, min_coeff=2e-3, damping=0.05, diag=diag)


@pytest.mark.parametrize("seed", range(4))
def test_term_budget_floor(seed):
    obs, layers, diag = random_case(300 + seed, 0, 12, n_layers=8, gates_per_layer=8)
    compare(0, 12, obs, layers, max_weight=3, min_coeff=1e-2, min_terms=40, diag=diag)


@pytest.mark.parametrize("defer", [True, False])
def test_clifford_frame_without_weight_cutoff(defer):
    obs, layers, diag = random_case(400, 0, 16, n_layers=8, gates_per_layer=8, clifford_frac=0.5)
    compare(0, 16, obs, layers, min_coeff=1e-3, diag=diag, defer=defer)


def test_kicked_ising_2d_first_steps():
    w = wl.config(2)
    compare(w.kind, w.n_units, w.observable, w.layers[:6], max_weight=w.max_weight, min_coeff=w.min_coeff,
            damping=w.damping, diag=w.diag_state)


def test_majorana_ucj_first_layers():
    w = wl.config(3)
    compare(w.kind, w.n_units, w.observable, w.layers[:40], max_weight=w.max_weight, min_coeff=w.min_coeff,
            damping=w.damping, diag=w.diag_state)


def test_hea_tfim_instance():
    w = wl.config(5)
    compare(w.kind, w.n_units, w.observable, w.layers[:12], max_weight=w.max_weight, min_coeff=w.min_coeff,
            diag=w.diag_state)


@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("kind,n_units", [(0, 12), (0, 60), (1, 20)])
def test_wide_generators_use_the_generic_child_sort(seed, kind, n_units):
    """Generators with more than 8 set bits: children are sorted by the LDS
    block sort + merge rounds instead of class buckets."""
    import numpy as np
    from tests.parity import random_angle
    rng = np.random.default_rng(500 + seed)
    def gen():
        k = 0
        for b in rng.choice(2 * n_units, size=int(rng.integers(9, 16)), replace=False):
            k |= 1 << int(b)
        return k
    layers = [[(gen(), random_angle(rng, 0.1, 0.05)) for _ in range(6)] for _ in range(6)]
    obs = [(gen(), 1.0), (gen(), 0.5)]
    compare(kind, n_units, obs, layers, max_weight=n_units // 2, min_coeff=1e-3, damping=0.02)


# Small stores run whole gate sequences in one workgroup (pq_set_small_runs);
# both paths must give the oracle's bits, and the same bits as each other.

@pytest.mark.parametrize("small", [False, True])
@pytest.mark.parametrize("seed", range(3))
def test_small_store_runs_and_chained_kernels(seed, small):
    obs, layers, diag = random_case(600 + seed, 0, 14, n_layers=6, gates_per_layer=8)
    compare(0, 14, obs, layers, max_weight=5, min_coeff=1e-3, damping=0.02, diag=diag, small_runs=small)


@pytest.mark.parametrize("kind,n_units", [(0, 20), (0, 50), (0, 100), (1, 24), (1, 40), (1, 100)])
def test_small_runs_hand_over_to_chained_kernels(kind, n_units):
    # lossless growth past the small-store limit inside one layer: the run
    # stops before the gate that could overflow it and the chained kernels go on
    import numpy as np
    from propaq_amd import runner
    from tests.parity import random_majorana_gen, random_pauli_gen
    rng = np.random.default_rng(610 + n_units)
    # generators on the first 8 units (shifted up in wide keys), so branches meet and the store grows
    shift = 2 * (n_units - 8) if kind == 0 else 0
    gen = ((lambda: random_pauli_gen(rng, 8) << shift) if kind == 0 else (lambda: random_majorana_gen(rng, 8)))
    layers = [[(gen(), float(rng.uniform(0.3, 1.2))) for _ in range(30)]]
    obs = [(gen(), 1.0), (gen(), 0.5)]
    d, o = compare(kind, n_units, obs, layers, diag=0, small_runs=True)
    assert max(o.n_terms) > 2048, max(o.n_terms)
    plain = runner.run_device(kind, n_units, obs, layers, want_terms=True, small_runs=False)
    assert plain.n_terms == d.n_terms
    assert plain.terms() == d.terms()
    for k in ("visited", "emitted", "declined", "hits", "created", "dropped_weight"):
        assert plain.counters[k] == d.counters[k], k


def test_small_runs_with_deferred_cliffords():
    # Clifford gates fold into the frame on the host; a run that stops keeps
    # only the folds of the gates it applied
    obs, layers, diag = random_case(620, 0, 12, n_layers=4, gates_per_layer=24, clifford_frac=0.5, phase_frac=0.1)
    compare(0, 12, obs, layers, min_coeff=1e-4, diag=diag, small_runs=True)
    compare(0, 12, obs, layers, max_weight=4, min_coeff=1e-4, diag=diag, small_runs=True)


@pytest.mark.parametrize("kind,n_units", [(0, 20), (1, 24)])
def test_small_run_stop_replays_deferred_cliffords(kind, n_units):
    # every third rotation is a Clifford that folds into the frame; the run
    # stops when the store outgrows LDS, and the host frame must then hold
    # exactly the folds of the gates before the stop
    import numpy as np
    from propaq_amd import runner
    from tests.parity import random_majorana_gen, random_pauli_gen
    rng = np.random.default_rng(630 + n_units)
    shift = 2 * (n_units - 8) if kind == 0 else 0
    gen = ((lambda: random_pauli_gen(rng, 8) << shift) if kind == 0 else (lambda: random_majorana_gen(rng, 8)))
    layer = []
    for i in range(45):
        layer.append((gen(), math.pi / 2 if i % 3 == 2 else float(rng.uniform(0.3, 1.2))))
    obs = [(gen(), 1.0), (gen(), 0.5)]
    d, o = compare(kind, n_units, obs, [layer], diag=0, small_runs=True)
    assert max(o.n_terms) > 8192, max(o.n_terms)
    plain = runner.run_device(kind, n_units, obs, [layer], want_terms=True, small_runs=False)
    assert plain.n_terms == d.n_terms and plain.terms() == d.terms()
    assert plain.counters["deferred"] == d.counters["deferred"] > 0
