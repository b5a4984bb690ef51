"""Host-side tests that need no GPU: the C ABI library loads and exports
every entry point its header declares, the drop-in datatypes / truncation /
plugin loader behave as the reference's do, and the synthetic workloads have
the advertised shapes."""

import math
import os
import re
import subprocess

import numpy as np
import pytest

import propaq_amd as pa
from oracle import oracle as orc
from propaq_amd import _native, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "propaq_amd.h")).read()
    declared = set(re.findall(r"\b(pq_[a-z_]+)\s*\(", header))
    assert declared and declared == set(_native.EXPORTED)
    lib = _native.load()
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.pq_abi_version() == 1


def test_engine_creation_fails_loudly_without_a_device():
    if _native.device_count() > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(RuntimeError, match="no HIP device"):
        _native.DeviceEngine(0, 20)


def test_pauli_string_key_and_algebra_match_oracle():
    rng = np.random.default_rng(1)
    for _ in range(200):
        a = pa.PauliString(int(rng.integers(0, 64)), int(rng.integers(0, 64)), 6)
        b = pa.PauliString(int(rng.integers(0, 64)), int(rng.integers(0, 64)), 6)
        assert pa.PauliString.from_key(a.key, 6) == a
        phase, prod = a @ b
        key, e = orc.product(orc.PAULI, a.key, b.key, 6)
        assert prod.key == key and phase == (1, 1j, -1, -1j)[e]
        assert a.commutes_with(b) == (not orc.anticommutes(orc.PAULI, a.key, b.key, 6))
        assert a.weight == orc.weight(orc.PAULI, a.key, 6)


def test_majorana_monomial_algebra_matches_oracle():
    rng = np.random.default_rng(2)
    for _ in range(200):
        a = pa.MajoranaMonomial(int(rng.integers(0, 1 << 12)), 12)
        b = pa.MajoranaMonomial(int(rng.integers(0, 1 << 12)), 12)
        phase, prod = a @ b
        key, e = orc.product(orc.MAJORANA, a.key, b.key, 6)
        assert prod.key == key and phase == (1, 1j, -1, -1j)[e]
        assert a.weight == orc.weight(orc.MAJORANA, a.key, 6)
        assert a.commutes_with(b) == (not orc.anticommutes(orc.MAJORANA, a.key, b.key, 6))
        for st in (0, 5, 63):
            assert a.trace_with_diag_state(st) == orc.trace(orc.MAJORANA, a.key, 6, st)
    with pytest.raises(ValueError):
        pa.MajoranaMonomial(1, 7)


def test_term_sum_ops_and_file_round_trip(tmp_path):
    ts = pa.PauliTermSum()
    ts.add(pa.PauliString(1, 0, 4), 0.5)
    ts.add(pa.PauliString(1, 0, 4), 0.25)
    ts[pa.PauliString(0, 3, 4)] = -1.0
    assert len(ts) == 2 and ts[pa.PauliString(1, 0, 4)] == 0.75
    assert abs(ts.norm_squared() - (0.75 ** 2 + 1.0)) < 1e-15
    p = tmp_path / "t.bin.gz"
    ts.save(str(p))
    back = pa.PauliTermSum.from_file(str(p))
    assert dict(back.items()) == dict(ts.items())
    m = pa.MajoranaTermSum()
    m.add(pa.MajoranaMonomial(0b1111, 8), 2.0)
    m.save(str(p))
    assert dict(pa.MajoranaTermSum.from_file(str(p)).items()) == dict(m.items())


def test_truncation_resolution():
    assert pa.resolve_truncation(None) == []
    ops = pa.resolve_truncation(pa.TruncationPolicy(weight_cutoff=3, coeff_cutoff=1e-3, min_terms=5))
    assert [type(o).__name__ for o in ops] == ["WeightTruncator", "CoefficientTruncator", "TermBudget"]
    assert pa.resolve_truncation(pa.TruncationPolicy(coeff_cutoff=0.0)) == []
    p = pa.TruncationPolicy(weight_cutoff=3, coeff_cutoff=0.5)
    assert not p.should_truncate(3, 0.5) and p.should_truncate(4, 0.5) and p.should_truncate(3, 0.4999)
    with pytest.raises(ValueError, match="surrogate"):
        pa.PauliPropagator(truncation=[pa.FrequencyTruncator(3)])
    with pytest.raises(TypeError):
        pa.PauliPropagator(truncation=3.0)
    prop = pa.PauliPropagator(truncation=pa.WeightTruncator(4))
    assert [t.weight for t in prop.truncators] == [4]


PLUGIN_SRC = r"""
#include <stdint.h>
#include <stddef.h>
#include <math.h>
#include <stdlib.h>
uint32_t propaq_noise_abi_version(void) { return ABI; }
#ifdef DEPENDS
uint32_t propaq_noise_depends(void) { return DEPENDS; }
#endif
#ifdef WITH_CREATE
void* propaq_noise_create(const char* j) { double* g = malloc(sizeof(double)); *g = 0.05; return g; }
#endif
#ifdef WITH_DESTROY
void propaq_noise_destroy(void* c) { free(c); }
#endif
double propaq_noise_factor(void* ctx, uint32_t kind, const uint64_t* w, size_t nw, uint32_t n, uint32_t weight,
                           uint32_t layer, uint32_t nl) {
  double g = ctx ? *(double*)ctx : 0.1;
  if (w) return (w[0] & 2) ? 0.5 : 1.0;
  return exp(-g * weight) * (1.0 - 0.01 * layer);
}
"""


def build_plugin(tmp_path, name, **defs):
    src = tmp_path / f"{name}.c"
    src.write_text(PLUGIN_SRC)
    so = tmp_path / f"{name}.so"
    flags = [f"-D{k}={v}" for k, v in defs.items()]
    subprocess.run(["gcc", "-shared", "-fPIC", "-O1", *flags, "-o", str(so), str(src), "-lm"], check=True)
    return str(so)


def test_native_noise_plugin_loading_rules(tmp_path):
    ok = pa.NativeNoiseModel(build_plugin(tmp_path, "ok", ABI=1))
    assert ok.depends == 0 and ok.abi_version == 1
    assert abs(ok.factor_term(0, [0], 4, 2) - math.exp(-0.2)) < 1e-15
    assert ok.table(0, 3) == [math.exp(-0.1 * w) for w in range(4)]
    key = pa.NativeNoiseModel(build_plugin(tmp_path, "key", ABI=1, DEPENDS=1))
    assert key.reads_key and key.factor_term(0, [2], 4, 1) == 0.5
    ctx = pa.NativeNoiseModel(build_plugin(tmp_path, "ctx", ABI=1, WITH_CREATE=1, WITH_DESTROY=1), "{}")
    assert abs(ctx.factor_term(0, [0], 4, 1) - math.exp(-0.05)) < 1e-15
    with pytest.raises(ValueError, match="ABI version"):
        pa.NativeNoiseModel(build_plugin(tmp_path, "abi2", ABI=2))
    with pytest.raises(ValueError, match="unknown depends"):
        pa.NativeNoiseModel(build_plugin(tmp_path, "dep", ABI=1, DEPENDS=4))
    with pytest.raises(ValueError, match="both"):
        pa.NativeNoiseModel(build_plugin(tmp_path, "half", ABI=1, WITH_CREATE=1))
    with pytest.raises(OSError):
        pa.NativeNoiseModel(str(tmp_path / "missing.so"))


def test_workload_shapes():
    assert workloads.config(1).n_gates == 390
    assert workloads.config(2).n_gates == 20 * (180 + 100)
    w3 = workloads.config(3)
    assert w3.n_gates == 4 * (496 * 2 + 496 * 3)
    rr = workloads.round_robin(32)
    assert len(rr) == 31 and len({p for r in rr for p in r}) == 496
    assert workloads.config(4).n_gates == 12 * (180 + 100)
    assert len(workloads.hea_batch(8)) == 8


def test_autotune_grid_and_tuned_defines():
    from propaq_amd import autotune, build_lib

    pts = list(autotune.points())
    n = 1
    for v in autotune.GRID.values():
        n *= len(v)
    assert len(pts) == n > 1 and len({autotune.name_of(p) for p in pts}) == n
    # every store tile fits a block's LDS (keys, value, flag and output slot per element, 160 KiB per CU)
    assert all(256 * p["PQ_STORE_ITEMS"] * (4 * 8 + 8 + 1 + 2) < 160 * 1024 for p in pts)
    # the child-merge tile must stay a power of two (the generic child sort is bitonic)
    assert all(p["PQ_MERGE_ITEMS"] & (p["PQ_MERGE_ITEMS"] - 1) == 0 for p in pts)
    d = build_lib.tuned_defines()
    assert all("=" in x and x.split("=")[0] in autotune.GRID for x in d)


def test_event_log_round_trip(tmp_path):
    # records written with the reference's keys and Rust float formatting read back unchanged
    from propaq_amd import events

    recs = [events.GateRecord(0, 0, None, 1, 3, 2, 0, 0.0, 0.0, 0.125),
            events.GateRecord(1, 0, 7, 3, 4, 1, 2, 1.5e-5, 1.0e-5, 2.5)]
    ph = events.phases_record(4, 2, {"branch": 3.0, "child_sort": 1.0, "store_merge": 2.0},
                              {"visited": 10, "emitted": 4, "declined": 2, "hits": 1})
    log = pa.Logger(str(tmp_path / "run.jsonl"), log_every=1)
    events.write_log(log, recs, 6, 1e-4, ph)
    lines = open(log.filename).read().splitlines()
    assert lines[0] == ('{"event":"gate","gate_idx":0,"layer_idx":0,"qiskit_gate_idx":null,"terms":3,'
                        '"ms_per_gate":1.250e-1}')
    assert '"discarded_coeff_l1":1.500000e-5' in lines[3] and '"coeff_cutoff":1.000000e-4' in lines[3]
    p = pa.LogParser(log.filename)
    assert p.terms == [3, 4] and p.gate_indices == [0, 1] and p.qiskit_gate_indices == [None, 7]
    assert p.terms_gained == [2, 1] and p.terms_discarded == [0, 2]
    assert p.discarded_coeff_l1 == [0.0, 1.5e-5] and p.discarded_coeff_max == [0.0, 1.0e-5]
    (e,) = p.engine_phases_events
    assert e.partitions == 1 and e.scan_s == 3e-3 and e.absorb_s == 3e-3 and e.emitted_share == 0.4
    assert e.scan_occupancy is None and e.inline_positions == 128
    events.write_log(pa.Logger(str(tmp_path / "b.jsonl"), log_every=2), recs, None, None, ph)
    assert pa.LogParser(str(tmp_path / "b.jsonl")).gate_indices == [0]
    with pytest.raises(ValueError):
        pa.Logger("x", log_every=0)
