"""propagate() and expectation_value(filename=) through the device (row f1,
crates/pauli/src/propagator.rs:157-261): the evolved term sum equals the
oracle's final terms with the deferred Clifford frame applied, bit for bit,
and the gzip term file (crates/core/src/interface/term_io.rs:13-89) written
by either call reads back to the same terms."""

import math

import pytest

import propaq_amd as pa
from oracle import oracle as orc
from tests.parity import random_case
from tests.test_gpu_noise_plugins import _objects

pytestmark = pytest.mark.gpu


CASES = [(0, 14, 950), (0, 40, 951), (0, 100, 952), (1, 12, 953), (1, 40, 954)]


@pytest.mark.parametrize("kind,n,seed", CASES)
@pytest.mark.parametrize("weight_cutoff", [None, 6])
def test_propagate_returns_the_oracle_terms(tmp_path, kind, n, seed, weight_cutoff):
    # without a weight cutoff the Clifford rotations fold into the frame, and
    # the export applies it (keys conjugated, coefficients signed)
    obs_terms, layers, diag = random_case(seed, kind, n, n_layers=6, gates_per_layer=12, clifford_frac=0.3)
    o = orc.run_oracle(kind, n, obs_terms, layers, max_weight=weight_cutoff, min_coeff=1e-3, diag_state=diag,
                       want_terms=True)
    prop_cls, obs, circ = _objects(kind, n, obs_terms, layers)
    prop = prop_cls(truncation=pa.TruncationPolicy(weight_cutoff=weight_cutoff, coeff_cutoff=1e-3))
    path = str(tmp_path / "terms.gz")
    evolved = prop.propagate(obs, circ, filename=path)
    assert dict(evolved._terms) == o.terms
    back = type(evolved).from_file(path)
    assert dict(back._terms) == o.terms
    # expectation_value writes the same file and reads the same value
    path2 = str(tmp_path / "terms2.gz")
    res = prop.expectation_value(obs, circ, initial_state=diag, filename=path2)
    assert res.n_terms == o.n_terms
    assert abs(res.expectation_value - o.expectation_value) <= 1e-10
    assert dict(type(evolved).from_file(path2)._terms) == o.terms


def test_propagate_large_kicked_ising_prefix():
    # 2D kicked Ising, six layers, damping: 6.8e4 terms through the sparse gates
    from propaq_amd import workloads as wl
    w = wl.config(2)
    o = orc.run_oracle(w.kind, w.n_units, w.observable, w.layers[:6], max_weight=w.max_weight,
                       min_coeff=w.min_coeff, noise_table=lambda li, nl: w.noise_table(), want_terms=True)
    prop_cls, obs, circ = _objects(w.kind, w.n_units, w.observable, w.layers[:6])
    prop = prop_cls(noise=pa.UniformNoiseModel(w.damping),
                    truncation=[pa.WeightTruncator(w.max_weight), pa.CoefficientTruncator(w.min_coeff)])
    evolved = prop.propagate(obs, circ)
    assert len(evolved) == len(o.terms) > 50_000
    assert dict(evolved._terms) == o.terms
    assert math.isclose(evolved.norm_squared(), sum(c * c for c in o.terms.values()), rel_tol=1e-12)
