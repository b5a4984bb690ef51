#!/usr/bin/env python
"""Throughput of the per-gate propagation path on one MI355X.

Workload (default): the 2D kicked-Ising circuit on an open 10 x 10 lattice
(100 qubits, 4 u64 words per key), 20 Trotter steps of RZZ(-pi/2) on all 180
edges then RX(0.6) on all 100 sites, observable Z on site (5,5), uniform
damping 1e-3 per layer, weight <= 12, |c| >= 1e-5, FP64 coefficients.

A step is one full propagation of that circuit (5600 gates, 40 noise layers)
from the seeded observable to the expectation value, with the term set
resident in HBM throughout.  value = term-gate updates / s, where a
term-gate update is one live term processed by one gate (sum over gates of
the live count at gate start).

roofline: gates over stores too large for the one-workgroup small-store runs
run as sparse in-place gates (k_sparse_gather + k_sparse_join); the unit is one
gate, whose algorithmic bytes are (N_in + N_out)(8W + 8) (one read and one
write of the store) and whose time is the two kernels' durations, stamped
with HIP events from their own dispatches on the engine's stream on every
--profile-every-th gate of the timed steps.  The sparse gate moves less than
that (only anticommuting rows are read in full and rewritten), so `achieved`
is an effective rate; `measured_GBps` is the HBM traffic per gate from the
rocprofv3 FETCH_SIZE / WRITE_SIZE summary under profiles/.  With --path
chained the store merge of the sorted chained kernels is reported instead.

cpu_baseline: the FP64 oracle (a C restatement of the reference's CPU
algorithm) on the first 12 layers of the same circuit, hash-partitioned over
one partition per thread on the process's CPUs (at most 16; `cores`, with
`nproc`, the CPU model and the one-core rate beside it), each repeated until
about two seconds have passed when it runs faster, with the device run of the
same 12 layers checked against it (live counts per gate).

adjoint: the differentiation workload (the same lattice, 12 steps, every RZZ
and RX with its own angle from U[-pi, pi], seed 4, 3360 angles, w <= 10,
|c| >= 1e-4, no noise): one plain forward run, then forward + reverse sweep
for all 3360 gradients with checkpoints every 4 layers (and a 64 GiB tape
budget), timed after one untimed recorded run (which allocates the tape
arena).  value = term-gate updates of the forward run / (T_fwd + T_adjoint).
The CPU comparison is parameter shift with the oracle: 2 forward runs per
angle, its forward rate timed on the first 12 layers applied.

batch: 256 seeded parameter sets of the 50-qubit hardware-efficient ansatz
with the TFIM observable in one batched store (see batch_leg), beside the
same instances run one at a time on the device and the oracle on a sample.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Measured HBM bytes of the store merge relative to its algorithmic bytes,
# from rocprofv3 FETCH_SIZE / WRITE_SIZE runs calibrated on a plain copy.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "store_merge_traffic.json")
SPARSE_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "sparse_gate_traffic.json")


def tuned_parameters():
    """Compile-time tile parameters of the library build (propaq_amd/tuned.json)."""
    from propaq_amd import build_lib

    return dict(x.split("=") for x in build_lib.tuned_defines())


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def prepare(eng, w, layers):
    """The run's inputs as arrays, built before any timing: the observable's
    keys and coefficients and, per layer in Heisenberg order (layers last
    first, gates last first), the generator words and angles."""
    from propaq_amd.runner import gates_to_arrays, keys_to_array
    import numpy as np

    W = eng.W
    obs = (keys_to_array([k for k, _ in w.observable], W), np.array([c for _, c in w.observable]))
    table = w.noise_table()
    heis = [list(reversed(l)) for l in reversed(layers)]
    if table is None:  # nothing happens between layers: one call applies them all
        heis = [[g for l in heis for g in l]]
    return obs, [gates_to_arrays(l, W) for l in heis], table


def run_once(eng, w, inputs, verbose=False):
    (keys, coeffs), heis, table = inputs
    eng.seed(keys, coeffs)
    n_terms = []
    t0 = time.perf_counter()
    for li, (gens, angles) in enumerate(heis):
        if table is not None:
            eng.scale_by_weight(table, reclaim=True)
        n_terms.extend(eng.apply_gates(gens, angles).tolist())
        if verbose:
            log(f"  layer {li + 1}/{len(heis)}: {n_terms[-1]} terms, {time.perf_counter() - t0:.2f} s")
    ev = eng.expectation(w.diag_state)
    return ev, n_terms


def adjoint_leg(args):
    """Forward and fwd+adjoint throughput on the differentiation workload."""
    import numpy as np

    from propaq_amd import _native, adjoint, runner, workloads

    w = workloads.config(4)
    eng = _native.DeviceEngine(w.kind, w.n_units)
    kw = dict(max_weight=w.max_weight, min_coeff=w.min_coeff, diag_state=w.diag_state)
    log("adjoint: plain forward run")
    runner.run_device(w.kind, w.n_units, w.observable, w.layers, engine=eng, **kw)  # warm-up
    eng.synchronize()
    eng.reset_counters()
    t0 = time.perf_counter()
    plain = runner.run_device(w.kind, w.n_units, w.observable, w.layers, engine=eng, **kw)
    eng.synchronize()
    t_plain = time.perf_counter() - t0
    updates = plain.counters["term_gate_updates"]
    log(f"adjoint: forward {t_plain:.3f} s; forward + reverse sweep (warm-up, then timed)")
    # the first recorded run also allocates the tape arena; the timed run reuses it
    adjoint.run_adjoint(w.kind, w.n_units, w.observable, w.layers, checkpoint_every=4, engine=eng, **kw)
    r = adjoint.run_adjoint(w.kind, w.n_units, w.observable, w.layers, checkpoint_every=4, engine=eng, **kw)
    n_par = len(r.gradient)
    log(f"adjoint: recorded forward {r.forward_s:.3f} s, sweep {r.backward_s:.3f} s, {r.checkpoints + 1} segments")
    out = {
        "workload": w.name, "parameters": n_par, "checkpoint_every_layers": 4, "segments": r.checkpoints + 1,
        "tape_bytes_peak": r.tape_bytes_peak,
        "value": updates / (r.forward_s + r.backward_s), "unit": "term-gate updates/s (fwd+adjoint)",
        "forward_value": updates / t_plain, "forward_s": t_plain, "recorded_forward_s": r.forward_s,
        "adjoint_s": r.backward_s, "adjoint_over_forward": r.backward_s / t_plain,
        "term_gate_updates": updates, "peak_terms": max(plain.n_terms), "expectation": r.expectation_value,
        "expectation_plain": plain.expectation_value, "gradient_norm": float(np.linalg.norm(r.gradient)),
    }
    if args.adjoint_check:
        # central differences of the device forward run at h = 1e-4 (the
        # workload's check) and 1e-7: a truncation decision flipped inside
        # [theta - h, theta + h] shows as a jump of about |c| ~ 1e-4 / 2h,
        # so only agreement at the small step pins the adjoint
        rng = np.random.default_rng(4)
        idx = rng.choice(n_par, size=args.adjoint_check, replace=False)
        flat = [(g, a) for layer in w.layers for g, a in layer]

        def value_at(j, d):
            f2 = list(flat)
            f2[j] = (f2[j][0], f2[j][1] + d)
            layers, pos = [], 0
            for layer in w.layers:
                layers.append(f2[pos:pos + len(layer)])
                pos += len(layer)
            return runner.run_device(w.kind, w.n_units, w.observable, layers, engine=eng, **kw).expectation_value

        rows = []
        for j in idx:
            fd = {h: (value_at(j, h) - value_at(j, -h)) / (2 * h) for h in (1e-4, 1e-7)}
            g = float(r.gradient[j])
            rows.append({"angle": int(j), "adjoint": g, "fd_1e-4": fd[1e-4], "fd_1e-7": fd[1e-7]})
            log(f"adjoint check {len(rows)}/{len(idx)}: angle {int(j)} adjoint {g:.6e} "
                f"fd(1e-4) {fd[1e-4]:.6e} fd(1e-7) {fd[1e-7]:.6e}")
        tol = lambda a, b: abs(a - b) <= 1e-6 * max(abs(a), abs(b), 1e-6)
        out["fd_check"] = {
            "angles": len(rows),
            "agree_h_1e-4": sum(tol(x["adjoint"], x["fd_1e-4"]) for x in rows),
            "agree_h_1e-7": sum(tol(x["adjoint"], x["fd_1e-7"]) for x in rows),
            "zero_gradient": sum(x["adjoint"] == 0.0 for x in rows),
            "max_abs_err_h_1e-7": max(abs(x["adjoint"] - x["fd_1e-7"]) for x in rows),
            "rows": rows}
    if not args.no_cpu:
        from oracle import oracle as orc

        # the first 12 layers in application (Heisenberg) order, all cores and one
        sample = w.layers[-12:]
        T = cpu_threads()
        rates = {}
        for threads in (T, 1):
            t1 = time.perf_counter()
            o = orc.run_oracle(w.kind, w.n_units, w.observable, sample, max_weight=w.max_weight,
                               min_coeff=w.min_coeff, diag_state=w.diag_state, defer=False,
                               **({"partitions": T, "threads": T} if threads > 1 else {}))
            rates[threads] = (o.term_gate_updates / (time.perf_counter() - t1), time.perf_counter() - t1)
        rate, cpu_s = rates[T]
        out["cpu_parameter_shift"] = {
            "value": rate / (2 * n_par), "unit": "term-gate updates/s (gradient by parameter shift)",
            "cores": T, "kind": "port", "forward_rate": rate, "one_core_forward_rate": rates[1][0],
            "one_core_value": rates[1][0] / (2 * n_par), **cpu_info(),
            "sample": f"oracle forward over the first 12 layers applied ({o.term_gate_updates} updates, "
                      f"{cpu_s:.2f} s on {T} threads, {rates[1][1]:.2f} s on one); parameter shift costs 2 "
                      f"forward runs per angle ({n_par} angles)"}
    eng.close()
    return out


def batch_leg(args):
    """Batched forward throughput: 256 seeded parameter sets of the 50-qubit
    hardware-efficient ansatz (RY, RZ, CZ ladders), TFIM observable, w <= 10,
    |c| >= 1e-5, all instances in one sorted store with one launch sequence
    per gate.  Depths are drawn from U{lo..hi} (default 1..5, which gives the
    per-instance term counts of 1e3 to 1e7 the workload is specified with;
    depths 4..16 reach 1e8 to 1e9 rows per instance at these cutoffs, far
    beyond one GPU).  Beside it: every instance run alone on the device
    (results must agree bit for bit) and the oracle on a bounded sample."""
    import numpy as np

    from propaq_amd import _native, batch, workloads
    from propaq_amd.runner import gates_to_arrays, keys_to_array

    n, B = 50, 256
    insts = workloads.hea_batch(B, seed=5, n=n, depth_lo=args.batch_depth_lo, depth_hi=args.batch_depth_hi)
    layers = [l for _, l in insts]
    obs = workloads.tfim_observable(n)
    kw = dict(max_weight=10, min_coeff=1e-5)
    eng = _native.DeviceEngine(workloads.PAULI, n, n_instances=B)
    log(f"batch: {B} instances, depths {args.batch_depth_lo}..{args.batch_depth_hi}: warm-up")
    batch.run_batch(workloads.PAULI, n, obs, layers, engine=eng, **kw)
    run = batch.run_batch(workloads.PAULI, n, obs, layers, engine=eng, **kw)
    eng.close()
    log(f"batch: {run.term_gate_updates:.3e} updates in {run.seconds:.3f} s, peak {max(run.n_terms)} rows")
    out = {"workload": f"hea-50q-tfim-batch{B}-depth{args.batch_depth_lo}-{args.batch_depth_hi}",
           "instances": B, "value": run.term_gate_updates / run.seconds, "unit": "term-gate updates/s (fwd)",
           "seconds": run.seconds, "term_gate_updates": run.term_gate_updates, "steps": run.steps,
           "peak_rows": int(max(run.n_terms)), "instance_rows_min": int(run.final_counts.min()),
           "instance_rows_max": int(run.final_counts.max())}
    # every instance alone on one device engine, inputs prepared beforehand
    solo = _native.DeviceEngine(workloads.PAULI, n)
    solo.set_cutoff(kw["max_weight"], kw["min_coeff"], None)
    W = solo.W
    ok = (keys_to_array([k for k, _ in obs], W), np.array([c for _, c in obs]))
    prepared = [[gates_to_arrays(list(reversed(l)), W) for l in reversed(ls)] for ls in layers]
    mismatch, t_solo, upd = 0, 0.0, 0
    for b, heis in enumerate(prepared):
        solo.reset_counters()
        t0 = time.perf_counter()
        solo.seed(*ok)
        for gens, angles in heis:
            solo.apply_gates(gens, angles)
        ev = solo.expectation(0)
        t_solo += time.perf_counter() - t0
        upd += solo.counters()["term_gate_updates"]
        mismatch += ev != run.expectation_values[b]
    solo.close()
    out["solo_device"] = {"value": upd / t_solo, "seconds": t_solo, "term_gate_updates": upd,
                          "expectation_mismatches": int(mismatch)}
    log(f"batch: solo device runs {t_solo:.3f} s, {mismatch} expectation mismatches")
    if not args.no_batch_grad:
        # forward + reverse sweep for every angle of every instance in one
        # batched run; depths up to batch_grad_depth_hi keep the tape and its
        # checkpoints well inside one GPU (the tape is 16 B per row per gate)
        ginst = workloads.hea_batch(B, seed=5, n=n, depth_lo=args.batch_depth_lo, depth_hi=args.batch_grad_depth_hi)
        glayers = [l for _, l in ginst]
        geng = _native.DeviceEngine(workloads.PAULI, n, n_instances=B)
        fwd = batch.run_batch(workloads.PAULI, n, obs, glayers, engine=geng, **kw)  # also the warm-up
        fwd = batch.run_batch(workloads.PAULI, n, obs, glayers, engine=geng, **kw)
        log(f"batch gradient: plain forward {fwd.seconds:.3f} s; forward + reverse sweep (warm-up, then timed)")
        batch.run_batch_gradient(workloads.PAULI, n, obs, glayers, engine=geng, checkpoint_every=4, **kw)
        gr = batch.run_batch_gradient(workloads.PAULI, n, obs, glayers, engine=geng, checkpoint_every=4, **kw)
        geng.close()
        n_par = int(sum(len(g) for g in gr.gradients))
        out["gradient"] = {
            "workload": f"hea-50q-tfim-batch{B}-depth{args.batch_depth_lo}-{args.batch_grad_depth_hi}",
            "parameters": n_par, "value": gr.term_gate_updates / (gr.forward_s + gr.backward_s),
            "unit": "term-gate updates/s (fwd+adjoint, all instances)", "forward_s": gr.forward_s,
            "adjoint_s": gr.backward_s, "plain_forward_s": fwd.seconds,
            "plain_forward_value": fwd.term_gate_updates / fwd.seconds,
            "adjoint_over_forward": gr.backward_s / fwd.seconds, "segments": gr.checkpoints + 1,
            "tape_bytes_peak": gr.tape_bytes_peak, "term_gate_updates": gr.term_gate_updates,
            "max_abs_value_diff_vs_plain_forward": float(np.max(np.abs(gr.expectation_values
                                                                        - fwd.expectation_values)))}
        log(f"batch gradient: {n_par} angles, recorded forward {gr.forward_s:.3f} s, sweep {gr.backward_s:.3f} s")
    if not args.no_cpu:
        from oracle import oracle as orc

        T = cpu_threads()
        t_cpu, cpu_upd, done = 0.0, 0, 0
        for b in range(B):
            t1 = time.perf_counter()
            o = orc.run_oracle(workloads.PAULI, n, obs, layers[b], partitions=T, threads=T, **kw)
            t_cpu += time.perf_counter() - t1
            cpu_upd += o.term_gate_updates
            done += 1
            if t_cpu > 10.0:
                break
        out["cpu_baseline"] = {"value": cpu_upd / t_cpu, "unit": "term-gate updates/s", "cores": T, "kind": "port",
                               "partitions": T, **cpu_info(),
                               "sample": f"oracle ({T} partitions on {T} threads) on the first {done} instances "
                                         f"one after another ({cpu_upd} updates, {t_cpu:.1f} s)"}
    return out


def surrogate_leg(args):
    """Surrogate evaluation throughput (crates/surrogate, evaluate_batch): a
    brickwork of symbolic X fields and XX / ZZ couplings on n qubits (one
    parameter per gate), observable mean Z, built once on the host without
    truncation, then evaluated on the device for a batch of random parameter
    sets.  Beside it: the oracle's tape evaluation on a bounded sample of the
    same sets (the first sets must agree bit for bit)."""
    import random

    import numpy as np

    from propaq_amd import _native, workloads
    from propaq_amd.runner import keys_to_array

    n, depth, n_sets = args.sur_qubits, args.sur_depth, args.sur_sets
    rng = random.Random(3)
    gates, p = [], 0
    for _ in range(depth):
        for q in range(n):
            gates.append((1 << 2 * q, p))
            p += 1
        for q in range(n - 1):
            zz = rng.random() >= 0.5
            gates.append(((2 << 2 * q) | (2 << 2 * (q + 1)) if zz else (1 << 2 * q) | (1 << 2 * (q + 1)), p))
            p += 1
    n_params = p
    t0 = time.perf_counter()
    b = _native.SurrogateBuilder(workloads.PAULI, n)
    b.seed(keys_to_array([2 << 2 * q for q in range(n)], b.W), np.full(n, 1.0 / n))
    heis = list(reversed(gates))
    counts = b.apply(keys_to_array([g for g, _ in heis], b.W), np.array([q for _, q in heis], dtype=np.int64),
                     np.zeros(len(heis)))
    model = b.compile(0, n_params)
    build_s = time.perf_counter() - t0
    b.close()
    log(f"surrogate: {n} qubits depth {depth}: {int(counts[-1])} terms, {model.n_terms} compiled, {model.n_ops} ops, "
        f"{model.n_levels} levels, built in {build_s:.2f} s")
    prng = np.random.default_rng(7)
    params = prng.uniform(-np.pi, np.pi, size=(n_sets, n_params))
    model.set_profiling(True)
    model.evaluate(params)  # warm-up
    reps, t_eval, tape_ms, sum_ms, tape_bytes, launches = 0, 0.0, 0.0, 0.0, 0.0, 0
    while reps < 3 or (t_eval < 2.0 and reps < 50):
        t1 = time.perf_counter()
        vals = model.evaluate(params)
        t_eval += time.perf_counter() - t1
        tm = model.timing()
        tape_ms += tm["tape_ms"]
        sum_ms += tm["sum_ms"]
        tape_bytes += tm["tape_bytes"]
        launches += tm["launches"]
        reps += 1
    # one parameter set per call, as an optimiser's inner loop calls evaluate
    model.set_profiling(False)
    one = params[:1]
    model.evaluate(one)
    t1 = time.perf_counter()
    for _ in range(20):
        model.evaluate(one)
    single_ms = (time.perf_counter() - t1) / 20 * 1e3
    model.set_profiling(True)
    achieved = tape_bytes / (tape_ms * 1e-3) / 1e9
    out = {"workload": f"surrogate-brickwork-{n}q-depth{depth}-{n_sets}sets", "unit": "parameter sets/s",
           "value": n_sets * reps / t_eval, "ops_per_s": model.n_ops * n_sets * reps / t_eval,
           "build_s": build_s, "live_terms": int(counts[-1]), "compiled_terms": model.n_terms, "ops": model.n_ops,
           "levels": model.n_levels, "monomials": model.n_monomials, "sets": n_sets, "reps": reps,
           "seconds_per_batch": t_eval / reps, "tape_ms_per_batch": tape_ms / reps, "sum_ms_per_batch": sum_ms / reps,
           "launches_per_batch": launches // reps, "single_set_evaluate_ms": single_ms,
           "roofline": {"bound": "hbm", "kernel": "k_tape_level_wide / k_tape_levels_block", "achieved": achieved,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                        "algorithmic_bytes_per_batch": tape_bytes / reps, "traffic": None}}
    tf = os.path.join(ROOT, "profiles", "surrogate_tape_traffic.json")
    if os.path.exists(tf):
        with open(tf) as fh:
            t = json.load(fh)
        if t.get("workload") == out["workload"]:
            out["roofline"]["traffic"] = t["measured_over_algorithmic"] * tape_bytes / reps
            out["roofline"]["traffic_source"] = t["source"]
    log(f"surrogate: {out['value']:.3e} sets/s, tape {tape_ms / reps:.2f} ms ({achieved:.0f} GB/s), "
        f"term sum {sum_ms / reps:.2f} ms; one set per call {single_ms:.3f} ms")
    if not args.no_cpu:
        from oracle import oracle as orc

        tape = model.export()
        roots = np.where(tape[6] == np.uint64(2 ** 64 - 1), -1, tape[6].astype(np.int64))
        done, t_cpu, same = 0, 0.0, True
        while t_cpu < 2.0 and done < n_sets:
            chunk = params[done:done + 16]
            t2 = time.perf_counter()
            ref = orc.tape_eval(*tape[:6], roots, n_params, chunk)
            t_cpu += time.perf_counter() - t2
            same = same and list(ref) == [float(v) for v in vals[done:done + len(chunk)]]
            done += len(chunk)
        # the last sets too: they come from the batch's last device pass
        tail = params[-16:]
        same_tail = list(orc.tape_eval(*tape[:6], roots, n_params, tail)) == [float(v) for v in vals[-16:]]
        out["cpu_baseline"] = {"value": done / t_cpu, "unit": "parameter sets/s", "cores": 1, "kind": "port",
                               "sample": f"oracle C tape evaluation (pqo_tape_eval) of the first {done} sets, "
                                         f"one after another ({t_cpu:.1f} s)",
                               "sets_bit_identical": bool(same), "last_16_sets_bit_identical": bool(same_tail)}
        log(f"surrogate: cpu {done / t_cpu:.3e} sets/s; {done} sets identical: {same}; last 16: {same_tail}")
    model.close()
    return out


def cpu_info():
    """The host the CPU baseline ran on: CPU model, nproc and the CPUs this
    process may use (its share on a shared box)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}


def cpu_threads():
    """Threads for the all-cores baseline: the CPUs this process may use, at
    most 16 (one GPU's share of a box)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def _timed_oracle(w, sample, **kw):
    """The oracle on `sample`, repeated until about two seconds have passed
    when one run is faster (so the rate is not a timer artefact)."""
    from oracle import oracle as orc

    table = w.noise_table()
    reps, secs, upd = 0, 0.0, 0
    while reps == 0 or (secs < 2.0 and reps < 1000):
        t1 = time.perf_counter()
        o = orc.run_oracle(w.kind, w.n_units, w.observable, sample, max_weight=w.max_weight, min_coeff=w.min_coeff,
                           noise_table=(lambda li, nl: table) if table else None, diag_state=w.diag_state, **kw)
        secs += time.perf_counter() - t1
        upd += o.term_gate_updates
        reps += 1
    return o, upd / secs, reps, secs


def cpu_baseline(w, sample):
    """The reference's CPU algorithm (the oracle, a C restatement of it) on a
    bounded sample: hash-partitioned over one partition per thread on all the
    process's cores (PartitionedTermSum), and on one core."""
    T = cpu_threads()
    o_all, rate_all, reps_all, s_all = _timed_oracle(w, sample, partitions=T, threads=T)
    o_one, rate_one, reps_one, s_one = _timed_oracle(w, sample)
    assert o_all.n_terms == o_one.n_terms, "partitioned and single-partition oracle runs differ"
    desc = f"first {len(sample)} layers of the same circuit ({o_one.term_gate_updates} term-gate updates)"
    return {"value": rate_all, "unit": "term-gate updates/s", "cores": T, "kind": "port",
            "partitions": T, "sample": f"{desc} x {reps_all} runs, {s_all:.1f} s",
            "one_core": {"value": rate_one, "cores": 1, "sample": f"{desc} x {reps_one} runs, {s_one:.1f} s"},
            **cpu_info(), "_counts": o_one.n_terms}


def _traffic_ratio(kernel, workload, layers):
    """Measured HBM bytes over algorithmic bytes for `kernel` on this workload,
    from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary under profiles/."""
    for fn in (TRAFFIC_FILE, SPARSE_TRAFFIC_FILE):
        if not os.path.exists(fn):
            continue
        with open(fn) as fh:
            t = json.load(fh)
        if t.get("kernel") == kernel and t.get("workload") == workload and t.get("layers") == layers:
            return t["measured_over_algorithmic"], t["source"]
    return None, None


def roofline_of(prof, w, n_layers):
    """The roofline line of the run's dominant gate path.

    Sparse gates (the default path): the unit is one gate, its device time the
    gather plus the join kernel, its algorithmic bytes (N_in + N_out)(8W + 8),
    one read and one write of the store.  The sparse design reads and writes
    less than that (only the anticommuting rows move), so `achieved` is an
    effective rate; `measured_GBps` is the HBM traffic the counters saw per
    gate over the same time.  Sorted chained kernels: the store merge, as
    before."""
    g, j, gate = prof["sparse_gather"], prof["sparse_join"], prof["sparse_gate"]
    if gate["launches"] > 0 and g["ms"] + j["ms"] > prof["store_merge"]["ms"]:
        n = gate["launches"]
        t_ms = (g["ms"] + j["ms"]) / n
        alg = gate["bytes"] / n
        achieved = alg / (t_ms * 1e-3) / 1e9
        kernel = "k_sparse_gather + k_sparse_join"
        ratio, src = _traffic_ratio(kernel, w.name, n_layers)
        traffic = ratio * alg if ratio is not None else None
        return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel,
                "unit_of_work": "one sparse gate (gather + join)", "launches": n, "avg_launch_ms": t_ms,
                "algorithmic_bytes_per_launch": alg, "algorithmic_bytes": "(N_in + N_out)(8W + 8) per gate",
                "measured_GBps": traffic / (t_ms * 1e-3) / 1e9 if traffic is not None else None,
                "measured_frac": traffic / (t_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic is not None else None,
                "traffic_source": src,
                "kernels": {"k_sparse_gather": {"avg_us": g["ms"] / n * 1e3, "design_bytes": g["bytes"] / n},
                            "k_sparse_join": {"avg_us": j["ms"] / n * 1e3, "design_bytes": j["bytes"] / n}}}
    m = prof["store_merge"]
    n = max(m["launches"], 1)
    avg_ms = m["ms"] / n
    alg = m["bytes"] / n
    achieved = alg / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    ratio, src = _traffic_ratio("k_store_merge", w.name, n_layers)
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": ratio * alg if ratio is not None else None,
            "kernel": "k_store_merge", "launches": m["launches"], "avg_launch_ms": avg_ms,
            "algorithmic_bytes_per_launch": alg, "traffic_source": src}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--layers", type=int, default=None, help="use only the first L circuit layers")
    ap.add_argument("--cpu-layers", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-every", type=int, default=16,
                    help="record phase events on every k-th gate (the store-merge roofline uses those launches)")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--path", choices=["default", "chained"], default="default",
                    help="default: small-store runs and sparse in-place gates; chained: sorted chained kernels only")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--sur-qubits", type=int, default=10)
    ap.add_argument("--sur-depth", type=int, default=4)
    ap.add_argument("--sur-sets", type=int, default=16384)
    ap.add_argument("--batch-depth-lo", type=int, default=1)
    ap.add_argument("--batch-depth-hi", type=int, default=5)
    ap.add_argument("--batch-grad-depth-hi", type=int, default=4)
    ap.add_argument("--no-batch-grad", action="store_true")
    ap.add_argument("--adjoint-check", type=int, default=0,
                    help="check this many seeded gradients against central differences (h=1e-4)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this engine runs on one GPU; the term store is not sharded across GPUs",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from propaq_amd import _native, workloads

    w = workloads.config(args.config)
    layers = w.layers if args.layers is None else w.layers[:args.layers]
    eng = _native.DeviceEngine(w.kind, w.n_units)
    eng.set_cutoff(w.max_weight, w.min_coeff, None)
    if args.path == "chained":
        eng.set_sparse(False)

    inputs = prepare(eng, w, layers)
    for i in range(args.warmup):
        log(f"warmup {i + 1}/{args.warmup}")
        run_once(eng, w, inputs, args.verbose)
    eng.synchronize()
    eng.reset_counters()
    eng.profile_enable(args.profile_every)
    log(f"timed: {args.steps} step(s)")
    t0 = time.perf_counter()
    evs, peak = [], 0
    for i in range(args.steps):
        ev, n_terms = run_once(eng, w, inputs, args.verbose)
        evs.append(ev)
        peak = max(peak, max(n_terms) if n_terms else 0)
        log(f"step {i + 1}/{args.steps}: <O> = {ev:.12f}, peak {peak} terms, {time.perf_counter() - t0:.2f} s")
    eng.synchronize()
    elapsed = time.perf_counter() - t0
    ctr = eng.counters()
    prof = eng.profile()
    host = eng.profile_host()
    updates = ctr["term_gate_updates"]
    value = updates / elapsed
    total_phase_ms = sum(p["ms"] for p in prof.values())
    roof = roofline_of(prof, w, len(layers))
    cpu = None
    if not args.no_cpu:
        sample = layers[:args.cpu_layers]
        log(f"cpu baseline: oracle on the first {len(sample)} layers, all cores and one core")
        cpu = cpu_baseline(w, sample)
        _, dev_counts = run_once(eng, w, prepare(eng, w, sample))
        cpu["device_counts_match"] = dev_counts == cpu.pop("_counts")
        log(f"cpu: {cpu['value']:.3e} updates/s on {cpu['cores']} threads, {cpu['one_core']['value']:.3e} on one; "
            f"device live counts match: {cpu['device_counts_match']}")

    adj = None if args.no_adjoint else adjoint_leg(args)
    bat = None if args.no_batch else batch_leg(args)
    sur = None if args.no_surrogate else surrogate_leg(args)

    out = {
        "metric": "term-gate updates/s (fwd) at fixed truncation; % HBM peak",
        "value": value,
        "unit": "term-gate updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / max(args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": w.name, "qubits": w.n_units, "key_words": eng.W,
                   "tile_parameters": tuned_parameters(),
                   "layers": len(layers), "gates": sum(len(l) for l in layers),
                   "max_weight": w.max_weight, "min_coeff": w.min_coeff, "damping": w.damping,
                   "peak_terms": peak, "term_gate_updates_per_step": updates // max(args.steps, 1),
                   "expectation": evs[-1] if evs else None},
        "roofline": roof,
        "phases_ms": {k: round(v["ms"], 3) for k, v in prof.items() if k != "sparse_gate"},
        "phases_share": {k: round(v["ms"] / total_phase_ms, 4) if total_phase_ms else 0.0 for k, v in prof.items()
                         if k != "sparse_gate"},
        "phases_note": "sums over the sampled gates (every --profile-every-th gate) and all noise passes of the "
                       f"{args.steps} timed steps, not per step",
        "host_ms": {"enqueue": round(host["enqueue_ms"], 3), "wait_for_counts": round(host["wait_ms"], 3),
                    "gates": host["gates"]},
        "cpu_baseline": cpu,
        "adjoint": adj,
        "batch": bat,
        "surrogate": sur,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
