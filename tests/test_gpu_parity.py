"""Device engine vs the FP64 oracle: term keys, signs and coefficients
bit-exact after every run, per-gate live counts identical, expectation within
1e-10 absolute.

Most cases run on four device paths: the sorted chained kernels alone
(branch count/scatter, class merge or generic sort, store merge), the sparse
in-place gates one at a time (gather, hash join with appended rows), the
sparse gates with consecutive commuting gates paired into one pass, and the
default (small-store runs in one workgroup while the store holds a few
thousand rows, sparse gates and pairs above).  The large cases reach 10^5 to
10^6 rows."""

import math

import pytest

from propaq_amd import workloads as wl
from tests.parity import compare, random_case

pytestmark = pytest.mark.gpu

PATH_OPTS = {"chained": dict(small_runs=False, sparse=False),
             "sparse": dict(small_runs=False, sparse=True, pairs=False),
             "pairs": dict(small_runs=False, sparse=True, pairs=True),
             "default": {}}
PATHS = pytest.mark.parametrize("path", list(PATH_OPTS))


def _path_ran(d, path, pairs_expected=True):
    c = d.counters
    if path == "chained":
        assert c["chained_gates"] > 0 and c["sparse_gates"] == 0 and c["small_gates"] == 0, c
    elif path == "sparse":
        assert c["sparse_gates"] > 0 and c["sparse_pairs"] == 0, c
        assert c["chained_gates"] == 0 and c["small_gates"] == 0, c
    elif path == "pairs":
        assert c["sparse_gates"] > 0 and c["chained_gates"] == 0 and c["small_gates"] == 0, c
        assert c["sparse_pairs"] > 0 or not pairs_expected, c
    else:
        assert c["chained_gates"] == 0, c


@PATHS
def test_kicked_ising_1d_full_run_bit_exact(path):
    w = wl.config(1)
    d, _ = compare(w.kind, w.n_units, w.observable, w.layers, max_weight=w.max_weight, min_coeff=w.min_coeff,
                   diag=w.diag_state, cache_key="cfg1", **PATH_OPTS[path])
    _path_ran(d, path)


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("n_units", [6, 30])
def test_random_pauli_lossless(seed, n_units):
    obs, layers, diag = random_case(seed, 0, n_units, n_layers=5, gates_per_layer=6)
    compare(0, n_units, obs, layers, diag=diag)


@PATHS
@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("n_units", [10, 40, 100, 200])
def test_random_pauli_truncated_noisy(seed, n_units, path):
    obs, layers, diag = random_case(100 + seed, 0, n_units, n_layers=8, gates_per_layer=10)
    compare(0, n_units, obs, layers, max_weight=4, min_coeff=2e-3, damping=0.05, diag=diag,
            cache_key=("rp", seed, n_units), **PATH_OPTS[path])


@PATHS
@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("n_units", [8, 24, 60, 100, 150])
def test_random_majorana_truncated_noisy(seed, n_units, path):
    obs, layers, diag = random_case(200 + seed, 1, n_units, n_layers=8, gates_per_layer=10)
    compare(1, n_units, obs, layers, max_weight=5, min_coeff=2e-3, damping=0.05, diag=diag,
            cache_key=("rm", seed, n_units), **PATH_OPTS[path])


def _commuting_layers(seed, n_units, n_layers=16, gates_per_layer=12):
    """Layers whose generators all commute (Z-type and X-type layers in turn,
    1 to 3 qubits each of a 14-qubit window at the top of the register, now
    and then a generator repeated in a row), so most consecutive sparse gates
    pair; Y on two window qubits as the observable."""
    import numpy as np
    from tests.parity import random_angle
    rng = np.random.default_rng(700 + seed)
    win = np.arange(max(0, n_units - 14), n_units)
    layers = []
    for li in range(n_layers):
        bit = 1 if li % 2 == 0 else 0  # Z = bit 2q+1, X = bit 2q
        layer = []
        for _ in range(gates_per_layer):
            if layer and rng.random() < 0.1:
                layer.append((layer[-1][0], random_angle(rng)))
                continue
            k = 0
            for q in rng.choice(win, size=int(rng.integers(1, 4)), replace=False):
                k |= 1 << (2 * int(q) + bit)
            layer.append((k, random_angle(rng)))
        layers.append(layer)
    obs = [(sum(3 << (2 * int(q)) for q in rng.choice(win, size=2, replace=False)), 1.0)]
    return obs, layers, int(rng.integers(0, 1 << min(n_units, 62)))


@pytest.mark.parametrize("path", ["sparse", "pairs"])
@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("n_units", [12, 40, 100, 220])
def test_commuting_layers_pair_up(seed, n_units, path):
    obs, layers, diag = _commuting_layers(seed, n_units)
    d, o = compare(0, n_units, obs, layers, max_weight=6, min_coeff=1e-4, damping=0.02, diag=diag,
                   cache_key=("cl", seed, n_units), **PATH_OPTS[path])
    assert max(o.n_terms) > 300
    _path_ran(d, path)
    if path == "pairs":  # the pair passes keep the counters of the gates one at a time
        from propaq_amd import runner
        single = runner.run_device(0, n_units, obs, layers, max_weight=6, min_coeff=1e-4,
                                   noise_table=lambda li, nl: runner.uniform_table(0.02, n_units), diag_state=diag,
                                   small_runs=False, sparse=True, pairs=False)
        for k in ("visited", "emitted", "declined", "hits", "created", "dropped_weight", "term_gate_updates"):
            assert single.counters[k] == d.counters[k], (k, single.counters[k], d.counters[k])
        assert math.isclose(single.counters["discarded_coeff_l1"], d.counters["discarded_coeff_l1"], rel_tol=1e-12)
        assert single.counters["discarded_coeff_max"] == d.counters["discarded_coeff_max"]


@pytest.mark.parametrize("seed", range(4))
def test_term_budget_floor(seed):
    obs, layers, diag = random_case(300 + seed, 0, 12, n_layers=8, gates_per_layer=8)
    compare(0, 12, obs, layers, max_weight=3, min_coeff=1e-2, min_terms=40, diag=diag)


@pytest.mark.parametrize("defer", [True, False])
def test_clifford_frame_without_weight_cutoff(defer):
    obs, layers, diag = random_case(400, 0, 16, n_layers=8, gates_per_layer=8, clifford_frac=0.5)
    compare(0, 16, obs, layers, min_coeff=1e-3, diag=diag, defer=defer)


# Workload prefixes.  The timed runs apply layers last-first (Heisenberg
# order), so the part of a run a prefix test covers is the circuit's last L
# layers, layers[-L:]; the kicked Ising lattice has no random angles, so its
# first and last layers are the same gates.

@PATHS
def test_kicked_ising_2d_first_steps(path):
    # W = 4, noise and reclaim every layer, 1.8 M rows at the end
    w = wl.config(2)
    d, o = compare(w.kind, w.n_units, w.observable, w.layers[-8:], max_weight=w.max_weight, min_coeff=w.min_coeff,
                   damping=w.damping, diag=w.diag_state, cache_key="cfg2-8", **PATH_OPTS[path])
    assert max(o.n_terms) > 1_000_000
    _path_ran(d, path)


@PATHS
def test_majorana_ucj_heisenberg_prefix(path):
    # W = 1 Majorana with the JW weight bound, coefficient cutoff and noise:
    # the timed run's first 80 layers, 1.25 M rows
    w = wl.config(3)
    d, o = compare(w.kind, w.n_units, w.observable, w.layers[-80:], max_weight=w.max_weight,
                   min_coeff=w.min_coeff, damping=w.damping, diag=w.diag_state, cache_key="cfg3-80", **PATH_OPTS[path])
    assert max(o.n_terms) > 1_000_000
    _path_ran(d, path)


def test_sparse_budget_hands_large_gates_to_chained_kernels():
    # lists and tables for about 3e5 rows at most: the store passes that on
    # its way to 1.25 M rows, so the largest gates run chained (the store
    # sorted first) and sparse gates resume when it shrinks below again
    from propaq_amd import _native

    w = wl.config(3)
    eng = _native.DeviceEngine(w.kind, w.n_units)
    eng.set_sparse_budget(40 << 20)
    d, o = compare(w.kind, w.n_units, w.observable, w.layers[-80:], max_weight=w.max_weight,
                   min_coeff=w.min_coeff, damping=w.damping, diag=w.diag_state, cache_key="cfg3-80", engine=eng)
    eng.close()
    c = d.counters
    assert c["sparse_gates"] > 0 and c["sparse_fallbacks"] > 0, c
    assert c["chained_gates"] == c["sparse_fallbacks"], c


@PATHS
def test_majorana_two_words_truncated_noisy(path):
    # W = 2 Majorana (40 sites): a two-repetition UCJ circuit's last 90 layers, 3.9e5 rows
    layers = wl.ucj_majorana(n_sites=40, reps=2, seed=7)[-90:]
    obs = [(0, 8.0)] + [((1 << (2 * p)) | (1 << (2 * p + 1)), 0.5) for p in range(16)]
    d, o = compare(1, 40, obs, layers, max_weight=6, min_coeff=1e-5, damping=1e-3, diag=(1 << 16) - 1,
                   cache_key="maj40", **PATH_OPTS[path])
    assert max(o.n_terms) > 300_000
    _path_ran(d, path)


@PATHS
def test_differentiation_workload_prefix(path):
    # W = 4, random angles, no noise: the last 6 layers, 9e4 rows
    w = wl.config(4)
    d, o = compare(w.kind, w.n_units, w.observable, w.layers[-6:], max_weight=w.max_weight, min_coeff=w.min_coeff,
                   diag=w.diag_state, cache_key="cfg4-6", **PATH_OPTS[path])
    assert max(o.n_terms) > 50_000
    _path_ran(d, path)


@PATHS
def test_hea_tfim_instance_noisy(path):
    # W = 2 Pauli with deferred CZ Cliffords and noise: one HEA instance's last 16 layers, 9.6e5 rows
    w = wl.config(5)
    d, o = compare(w.kind, w.n_units, w.observable, w.layers[-16:], max_weight=w.max_weight, min_coeff=w.min_coeff,
                   damping=1e-3, diag=w.diag_state, cache_key="cfg5-16", **PATH_OPTS[path])
    assert max(o.n_terms) > 500_000
    _path_ran(d, path)


def test_hea_tfim_instance():
    w = wl.config(5)
    compare(w.kind, w.n_units, w.observable, w.layers[-12:], max_weight=w.max_weight, min_coeff=w.min_coeff,
            diag=w.diag_state)


def _wide_case(seed, kind, n_units):
    """Generators with 9 to 14 set bits (more than the 8 class bits), a weight
    bound the observable passes, and a store that grows to thousands of rows."""
    import numpy as np
    from tests.parity import random_angle
    rng = np.random.default_rng(500 + seed)
    span = min(2 * n_units, 24)  # bits kept near the bottom so branches meet again

    def gen(lo, hi):
        k = 0
        for b in rng.choice(span, size=int(rng.integers(lo, hi)), replace=False):
            k |= 1 << int(b)
        return k
    layers = [[(gen(9, 15), random_angle(rng, 0.05, 0.02)) for _ in range(8)] for _ in range(4)]
    obs = [(gen(2, 4), 1.0), (gen(2, 4), 0.5)]
    return obs, layers


@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("kind,n_units", [(0, 12), (0, 60), (1, 20), (1, 40)])
def test_wide_generators_use_the_generic_child_sort(seed, kind, n_units):
    """Children of generators with more than 8 set bits are sorted by the LDS
    block sort + merge rounds instead of class buckets."""
    obs, layers = _wide_case(seed, kind, n_units)
    d, o = compare(kind, n_units, obs, layers, max_weight=n_units, min_coeff=1e-3, damping=0.01, small_runs=False,
                   sparse=False, cache_key=("wide", seed, kind, n_units))
    assert d.counters["generic_sorts"] > 0
    assert max(o.n_terms) > 2000, max(o.n_terms)
    # the sparse path takes any generator
    d, o = compare(kind, n_units, obs, layers, max_weight=n_units, min_coeff=1e-3, damping=0.01, small_runs=False,
                   cache_key=("wide", seed, kind, n_units))
    assert d.counters["sparse_gates"] > 0


# Small stores run whole gate sequences in one workgroup (pq_set_small_runs);
# both paths must give the oracle's bits, and the same bits as each other.

@PATHS
@pytest.mark.parametrize("seed", range(3))
def test_small_store_runs_and_chained_kernels(seed, path):
    obs, layers, diag = random_case(600 + seed, 0, 14, n_layers=6, gates_per_layer=8)
    compare(0, 14, obs, layers, max_weight=5, min_coeff=1e-3, damping=0.02, diag=diag, cache_key=("ss", seed),
            **PATH_OPTS[path])


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
    for sparse in (False, True):
        plain = runner.run_device(kind, n_units, obs, layers, want_terms=True, small_runs=False, sparse=sparse)
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
    for sparse in (False, True):
        plain = runner.run_device(kind, n_units, obs, [layer], want_terms=True, small_runs=False, sparse=sparse)
        assert plain.n_terms == d.n_terms and plain.terms() == d.terms()
        assert plain.counters["deferred"] == d.counters["deferred"] > 0
