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

majorana: the Majorana UCJ workload (32 sites, 64 modes, W = 1, 4
repetitions of a Givens network plus density-density phases, JW weight <= 6,
|c| >= 1e-5, damping 1e-3 per layer), whose stores stay mid-size (1e5 to
3.4e6 rows), measured the same way, with the oracle on its first 96 layers.

adjoint: the differentiation workload (the same lattice, 12 steps, every RZZ
and RX with its own angle from U[-pi, pi], seed 4, 3360 angles, w <= 10,
|c| >= 1e-4, no noise): one plain forward run, then forward + reverse sweep
for all 3360 gradients with checkpoints every 4 layers (and a 64 GiB tape
budget), timed after one untimed recorded run (which allocates the tape
arena).  value = term-gate updates of the forward run / (T_fwd + T_adjoint).
Then 32 seeded angles with nonzero gradient against central differences of
the device forward run at h = 1e-4 and 1e-7, 4 against parameter shift on
the device, and 8 of the first 4 steps against parameter shift with the
oracle.  The CPU comparison is parameter shift with the oracle: 2 forward
runs per angle, its forward rate timed on the first 12 layers applied.

batch: 256 seeded parameter sets of the 50-qubit hardware-efficient ansatz
with the TFIM observable in one batched store at depths U{1..5} (a secondary
workload where every instance fits), beside the same instances run one at a
time on the device and the oracle on a sample; batch_as_specified: the
depth-U{4..16} draw on a bounded sample (see batch_spec_leg).
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


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


def _seeded_angles(gradient, count, seed=4):
    """`count` angles drawn with a seeded generator from those whose adjoint
    gradient is nonzero (most of the circuit's angles lie outside the
    observable's light cone or reach it only through truncated terms: their
    gradient is exactly zero by either method), plus how many are nonzero."""
    import numpy as np

    nonzero = np.flatnonzero(np.asarray(gradient) != 0.0)
    rng = np.random.default_rng(seed)
    pick = rng.choice(nonzero, size=min(count, len(nonzero)), replace=False) if len(nonzero) else []
    return sorted(int(j) for j in pick), int(len(nonzero))


def _shifted_layers(layers, j, d):
    """`layers` with the angle of the j-th rotation (circuit order) moved by d."""
    out, pos = [], 0
    for layer in layers:
        out.append([(g, a + d) if pos + i == j else (g, a) for i, (g, a) in enumerate(layer)])
        pos += len(layer)
    return out


def gradient_checks(args, w, gradient, kw):
    """The adjoint gradients of --adjoint-check seeded angles (drawn from the
    nonzero ones) against central differences of the device forward run at
    h = 1e-4 and h = 1e-7 (a truncation decision flipped inside [theta - h,
    theta + h] shows as a jump of about |c| / 2h, so where h = 1e-4 disagrees
    only the small step pins the angle), and against parameter shift, (f(theta
    + pi/2) - f(theta - pi/2)) / 2, exact for exp(-i theta P / 2) without
    truncation, on the device for --shift-angles of them; then the oracle
    parameter-shift check on a prefix of the circuit.  The forward runs use a
    plain engine with Clifford deferral off, the function the adjoint
    differentiates."""
    import math

    from propaq_amd import _native, runner

    idx, nonzero = _seeded_angles(gradient, args.adjoint_check)
    eng = _native.DeviceEngine(w.kind, w.n_units)

    def value_at(j, d, layers=w.layers):
        return runner.run_device(w.kind, w.n_units, w.observable, _shifted_layers(layers, j, d), engine=eng,
                                 defer=False, **kw).expectation_value

    # relative 1e-6, plus the rounding of the expectation's sum (about 1e-17
    # here, |<O>| ~ 2e-3) divided by the step
    tol = lambda a, b, h: abs(a - b) <= 1e-6 * max(abs(a), abs(b)) + 1e-17 / h
    rows = []
    t0 = time.perf_counter()
    for j in idx:
        g = float(gradient[j])
        row = {"angle": j, "adjoint": g}
        for h in (1e-4, 1e-7):
            row[f"fd_{h:.0e}"] = (value_at(j, h) - value_at(j, -h)) / (2 * h)
        if not (tol(g, row["fd_1e-04"], 1e-4) or tol(g, row["fd_1e-07"], 1e-7)):
            # a decision flips within 1e-7 of theta: it lies on one side, so
            # one of the one-sided differences is free of it
            f0, fp, fm = value_at(j, 0.0), value_at(j, 1e-7), value_at(j, -1e-7)
            row["one_sided_1e-07"] = min(((fp - f0) / 1e-7, (f0 - fm) / 1e-7), key=lambda v: abs(v - g))
        rows.append(row)
        log(f"adjoint check {len(rows)}/{len(idx)}: angle {j} adjoint {g:.6e} fd(1e-4) {row['fd_1e-04']:.6e} "
            f"fd(1e-7) {row['fd_1e-07']:.6e}"
            + (f" one-sided(1e-7) {row['one_sided_1e-07']:.6e}" if "one_sided_1e-07" in row else ""))
    one_ok = lambda x: "one_sided_1e-07" in x and abs(x["adjoint"] - x["one_sided_1e-07"]) <= 1e-5 * abs(x["adjoint"])
    fd = {"angles": len(rows), "seed": 4, "nonzero_gradients": nonzero, "parameters": len(gradient),
          "seconds": time.perf_counter() - t0,
          "agree_h_1e-4": sum(tol(x["adjoint"], x["fd_1e-04"], 1e-4) for x in rows),
          "agree_h_1e-7": sum(tol(x["adjoint"], x["fd_1e-07"], 1e-7) for x in rows),
          "agree_either": sum(tol(x["adjoint"], x["fd_1e-04"], 1e-4) or tol(x["adjoint"], x["fd_1e-07"], 1e-7)
                              for x in rows),
          "one_sided_checked": sum("one_sided_1e-07" in x for x in rows),
          "one_sided_agree": sum(one_ok(x) for x in rows),
          "max_rel_err_h_1e-7": max((abs(x["adjoint"] - x["fd_1e-07"]) / max(abs(x["adjoint"]), 1e-300)
                                     for x in rows), default=None),
          "tolerance": "central: |adjoint - fd| <= 1e-6 max(|adjoint|, |fd|) + 1e-17 / h; one-sided (only where "
                       "neither central difference agrees): relative 1e-5", "rows": rows}
    ps_rows = []
    half = math.pi / 2
    for j in idx[:args.shift_angles]:
        lo, hi = value_at(j, -half), value_at(j, half)
        ps_rows.append({"angle": j, "adjoint": float(gradient[j]), "device_shift": (hi - lo) / 2})
        log(f"parameter shift {len(ps_rows)}: angle {j} adjoint {gradient[j]:.6e} device {ps_rows[-1]['device_shift']:.6e}")
    ps = {"angles": len(ps_rows), "rows": ps_rows,
          "max_abs_adjoint_minus_shift": max((abs(x["adjoint"] - x["device_shift"]) for x in ps_rows), default=None),
          "note": "parameter shift differentiates the truncated forward run at shifted angles, where other terms "
                  "fall under the cutoffs, so it agrees with the adjoint only to the truncation error"}
    if args.shift_oracle_steps and not args.no_cpu:
        ps["oracle_prefix"] = oracle_shift_check(args, w, eng, kw)
    eng.close()
    return fd, ps


def oracle_shift_check(args, w, eng, kw):
    """Parameter shift with the oracle on the circuit's first
    --shift-oracle-steps steps (two oracle forward runs per angle, angles
    drawn from those whose adjoint gradient on that prefix is nonzero): the
    device's forward values at the shifted angles must equal the oracle's
    within 1e-10, and the device adjoint of the prefix is set beside both
    shift gradients."""
    import math

    from oracle import oracle as orc
    from propaq_amd import _native, adjoint, runner

    prefix = w.layers[:2 * args.shift_oracle_steps]
    aeng = _native.DeviceEngine(w.kind, w.n_units)
    grad = adjoint.run_adjoint(w.kind, w.n_units, w.observable, prefix, checkpoint_every=4, engine=aeng,
                               **kw).gradient
    aeng.close()
    idx, nonzero = _seeded_angles(grad, args.shift_oracle_angles)
    T = cpu_threads()
    rows, t0 = [], time.perf_counter()
    for j in idx:
        vals = {}
        for d in (-math.pi / 2, math.pi / 2):
            layers = _shifted_layers(prefix, j, d)
            dev = runner.run_device(w.kind, w.n_units, w.observable, layers, engine=eng, defer=False,
                                    **kw).expectation_value
            ora = orc.run_oracle(w.kind, w.n_units, w.observable, layers, max_weight=w.max_weight,
                                 min_coeff=w.min_coeff, diag_state=w.diag_state, defer=False, partitions=T,
                                 threads=T).expectation_value
            vals[d] = (dev, ora)
        (dl, ol), (dh, oh) = vals[-math.pi / 2], vals[math.pi / 2]
        rows.append({"angle": j, "adjoint": float(grad[j]), "device_shift": (dh - dl) / 2, "oracle_shift": (oh - ol) / 2,
                     "max_abs_forward_diff": max(abs(dl - ol), abs(dh - oh))})
        log(f"oracle parameter shift {len(rows)}/{len(idx)}: angle {j} adjoint {grad[j]:.6e} "
            f"oracle {rows[-1]['oracle_shift']:.6e} forward diff {rows[-1]['max_abs_forward_diff']:.1e}")
    return {"steps": args.shift_oracle_steps, "angles": len(rows), "nonzero_gradients": nonzero,
            "parameters": len(grad), "seconds": time.perf_counter() - t0, "oracle_threads": T,
            "forward_within_1e-10": sum(x["max_abs_forward_diff"] <= 1e-10 for x in rows),
            "device_shift_equals_oracle_shift_1e-10": sum(abs(x["device_shift"] - x["oracle_shift"]) <= 1e-10
                                                          for x in rows),
            "max_abs_adjoint_minus_oracle_shift": max((abs(x["adjoint"] - x["oracle_shift"]) for x in rows),
                                                      default=None), "rows": rows}


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
    ck = args.adjoint_checkpoint_every or None
    tk = dict(checkpoint_every=ck, tape_budget=int(args.adjoint_tape_gib * (1 << 30)), engine=eng)
    adjoint.run_adjoint(w.kind, w.n_units, w.observable, w.layers, **tk, **kw)
    r = adjoint.run_adjoint(w.kind, w.n_units, w.observable, w.layers, **tk, **kw)
    n_par = len(r.gradient)
    log(f"adjoint: recorded forward {r.forward_s:.3f} s, sweep {r.backward_s:.3f} s, {r.checkpoints + 1} segments")
    out = {
        "workload": w.name, "parameters": n_par, "checkpoint_every_layers": ck,
        "tape_budget_bytes": tk["tape_budget"], "segments": r.checkpoints + 1,
        "tape_bytes_peak": r.tape_bytes_peak,
        "value": updates / (r.forward_s + r.backward_s), "unit": "term-gate updates/s (fwd+adjoint)",
        "forward_value": updates / t_plain, "forward_s": t_plain, "recorded_forward_s": r.forward_s,
        "adjoint_s": r.backward_s, "adjoint_over_forward": r.backward_s / t_plain,
        "term_gate_updates": updates, "peak_terms": max(plain.n_terms), "expectation": r.expectation_value,
        "expectation_plain": plain.expectation_value, "gradient_norm": float(np.linalg.norm(r.gradient)),
    }
    if ck is None:  # beside it the memory-lean schedule: checkpoints every 4 layers, earlier segments replayed
        c4 = adjoint.run_adjoint(w.kind, w.n_units, w.observable, w.layers, checkpoint_every=4,
                                 tape_budget=tk["tape_budget"], engine=eng, **kw)
        out["checkpointed_every_4_layers"] = {
            "value": updates / (c4.forward_s + c4.backward_s), "recorded_forward_s": c4.forward_s,
            "adjoint_s": c4.backward_s, "segments": c4.checkpoints + 1, "tape_bytes_peak": c4.tape_bytes_peak,
            "max_abs_gradient_diff": float(np.max(np.abs(np.asarray(c4.gradient) - np.asarray(r.gradient))))}
        log(f"adjoint: checkpointed every 4 layers: recorded forward {c4.forward_s:.3f} s, sweep "
            f"{c4.backward_s:.3f} s, {c4.checkpoints + 1} segments")
    if args.adjoint_check:
        out["fd_check"], out["parameter_shift"] = gradient_checks(args, w, r.gradient, kw)
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
           "role": "secondary: a shallower depth draw whose 256 instances all fit one GPU; the specified "
                   "depth U{4..16} is batch_as_specified",
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
        # checkpoints well inside one GPU (the tape is 16-20 B per row a gate
        # changes)
        ginst = workloads.hea_batch(B, seed=5, n=n, depth_lo=args.batch_depth_lo, depth_hi=args.batch_grad_depth_hi)
        glayers = [l for _, l in ginst]
        geng = _native.DeviceEngine(workloads.PAULI, n, n_instances=B)
        fwd = batch.run_batch(workloads.PAULI, n, obs, glayers, engine=geng, **kw)  # also the warm-up
        fwd = batch.run_batch(workloads.PAULI, n, obs, glayers, engine=geng, **kw)
        log(f"batch gradient: plain forward {fwd.seconds:.3f} s; forward + reverse sweep (warm-up, then timed)")
        gk = dict(checkpoint_every=args.adjoint_checkpoint_every or None,
                  tape_budget=int(args.adjoint_tape_gib * (1 << 30)))
        batch.run_batch_gradient(workloads.PAULI, n, obs, glayers, engine=geng, **gk, **kw)
        geng.reset_counters()
        gr = batch.run_batch_gradient(workloads.PAULI, n, obs, glayers, engine=geng, **gk, **kw)
        gc = geng.counters()
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
                                                                        - fwd.expectation_values))),
            "paths": {k: int(gc[k]) for k in ("sparse_gates", "chained_gates", "sparse_fallbacks")}}
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


def batch_spec_leg(args):
    """The batching workload as specified: 256 seeded parameter sets of the
    50-qubit hardware-efficient ansatz with depth from U{4..16} (seed 5), TFIM
    observable, w <= 10, |c| >= 1e-5.  Per-instance peaks grow about 3-5x per
    unit of depth at these cutoffs (0.9 M, 4.7 M, 24 M and 76 M rows at depths
    4 to 7 measured here; more than 1.2e8 at depth 8, so some 1e9 at depth 10),
    so most of the draw does not fit one GPU.  Bounded sample: of the first
    --spec-instances instances of the draw, those with depth <= --spec-max-depth
    (9: the deepest whose instances fit beside the others) run as one batched store
    (their predicted peaks sum to about 2e8 rows); the deeper ones are listed
    as not run.  Beside it: the oracle on the sample's instances, smallest
    first, for about ten seconds."""
    import collections

    import numpy as np

    from propaq_amd import _native, batch, workloads

    n = 50
    draw = workloads.hea_batch(256, seed=5, n=n, depth_lo=4, depth_hi=16)
    head = draw[:args.spec_instances]
    pick = [b for b, (d, _) in enumerate(head) if d <= args.spec_max_depth]
    layers = [head[b][1] for b in pick]
    obs = workloads.tfim_observable(n)
    kw = dict(max_weight=10, min_coeff=1e-5)
    out = {"workload": "hea-50q-tfim-batch256-depth4-16", "draw_depths": dict(sorted(collections.Counter(
               d for d, _ in draw).items())), "sample_of_first": args.spec_instances,
           "instances": len(pick), "instance_indices": pick, "instance_depths": [head[b][0] for b in pick],
           "not_run": {"indices": [b for b in range(len(head)) if b not in pick],
                       "reason": f"depth > {args.spec_max_depth}: predicted peak of 1e9 rows or more per instance"}}
    if not pick:
        return out
    eng = _native.DeviceEngine(workloads.PAULI, n, n_instances=len(pick))
    log(f"batch (as specified): {len(pick)} of the first {len(head)} instances, depths "
        f"{sorted(head[b][0] for b in pick)}: warm-up")
    batch.run_batch(workloads.PAULI, n, obs, layers, engine=eng, **kw)
    run = batch.run_batch(workloads.PAULI, n, obs, layers, engine=eng, **kw)
    eng.close()
    log(f"batch (as specified): {run.term_gate_updates:.3e} updates in {run.seconds:.3f} s, "
        f"peak {max(run.n_terms)} rows")
    out.update({"value": run.term_gate_updates / run.seconds, "unit": "term-gate updates/s (fwd)",
                "seconds": run.seconds, "term_gate_updates": run.term_gate_updates, "steps": run.steps,
                "peak_rows": int(max(run.n_terms)), "instance_rows_final": [int(v) for v in run.final_counts],
                "expectation_values": [float(v) for v in run.expectation_values]})
    if not args.no_cpu:
        from oracle import oracle as orc

        T = cpu_threads()
        order = sorted(range(len(pick)), key=lambda i: (head[pick[i]][0], i))
        t_cpu, cpu_upd, done, worst = 0.0, 0, [], 0.0
        for i in order:
            t1 = time.perf_counter()
            o = orc.run_oracle(workloads.PAULI, n, obs, layers[i], partitions=T, threads=T, **kw)
            t_cpu += time.perf_counter() - t1
            cpu_upd += o.term_gate_updates
            done.append(pick[i])
            worst = max(worst, abs(o.expectation_value - run.expectation_values[i]))
            if t_cpu > 10.0:
                break
        out["cpu_baseline"] = {"value": cpu_upd / t_cpu, "unit": "term-gate updates/s", "cores": T, "kind": "port",
                               "partitions": T, **cpu_info(), "max_abs_expectation_diff_vs_device": worst,
                               "sample": f"oracle ({T} partitions on {T} threads) on instances {done} of the "
                                         f"sample, smallest first ({cpu_upd} updates, {t_cpu:.1f} s)"}
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
    from the rocprofv3 FETCH_SIZE / WRITE_SIZE summaries under profiles/
    (profiles/*_traffic.json)."""
    import glob

    for fn in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_traffic.json"))):
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
        kernel = "k_sparse_gather + k_sparse_join (k_pair_gather + k_pair_join for gate pairs)"
        ratio, src = _traffic_ratio(kernel, w.name, n_layers)
        traffic = ratio * alg if ratio is not None else None
        return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel,
                "unit_of_work": "one sparse gate (a gather and a join per gate, or per two commuting gates)",
                "launches": n, "avg_launch_ms": t_ms,
                "algorithmic_bytes_per_launch": alg, "algorithmic_bytes": "(N_in + N_out)(8W + 8) per gate",
                "measured_GBps": traffic / (t_ms * 1e-3) / 1e9 if traffic is not None else None,
                "measured_frac": traffic / (t_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic is not None else None,
                "traffic_source": src,
                "kernels": {"gather": {"avg_us": g["ms"] / n * 1e3, "design_bytes": g["bytes"] / n},
                            "join": {"avg_us": j["ms"] / n * 1e3, "design_bytes": j["bytes"] / n}},
                "per_gate_note": "kernel time and design bytes per gate; a pair pass's are split over its two gates "
                                 "(its design bytes are counted for single gates only)"}
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


def measure(w, layers, steps, warmup, profile_every, path, cpu_layers, verbose=False):
    """Warm-up, then `steps` timed propagations of `layers` with the inputs
    resident on the device; the roofline of the dominant gate path from the
    profiled gates; the CPU baseline on the first `cpu_layers` layers (none
    when cpu_layers is None), with the device's live counts checked against it."""
    from propaq_amd import _native

    eng = _native.DeviceEngine(w.kind, w.n_units)
    eng.set_cutoff(w.max_weight, w.min_coeff, None)
    if path == "chained":
        eng.set_sparse(False)
    inputs = prepare(eng, w, layers)
    for i in range(warmup):
        log(f"{w.name}: warmup {i + 1}/{warmup}")
        run_once(eng, w, inputs, verbose)
    eng.synchronize()
    eng.reset_counters()
    eng.profile_enable(profile_every)
    log(f"{w.name}: timed {steps} step(s)")
    t0 = time.perf_counter()
    evs, peak = [], 0
    for i in range(steps):
        ev, n_terms = run_once(eng, w, inputs, verbose)
        evs.append(ev)
        peak = max(peak, max(n_terms) if n_terms else 0)
        log(f"{w.name}: step {i + 1}/{steps}: <O> = {ev:.12f}, peak {peak} terms, {time.perf_counter() - t0:.2f} s")
    eng.synchronize()
    elapsed = time.perf_counter() - t0
    ctr = eng.counters()
    prof = eng.profile()
    host = eng.profile_host()
    updates = ctr["term_gate_updates"]
    total_phase_ms = sum(p["ms"] for p in prof.values())
    cpu = None
    if cpu_layers is not None:
        sample = layers[:cpu_layers]
        log(f"{w.name}: cpu baseline, oracle on the first {len(sample)} layers, all cores and one core")
        cpu = cpu_baseline(w, sample)
        eng.profile_enable(0)
        _, dev_counts = run_once(eng, w, prepare(eng, w, sample))
        cpu["device_counts_match"] = dev_counts == cpu.pop("_counts")
        log(f"{w.name}: cpu {cpu['value']:.3e} updates/s on {cpu['cores']} threads, {cpu['one_core']['value']:.3e} "
            f"on one; device live counts match: {cpu['device_counts_match']}")
    out = {
        "value": updates / elapsed,
        "ms_per_step": elapsed * 1e3 / max(steps, 1),
        "config": {"workload": w.name, "units": w.n_units, "key_words": eng.W,
                   "tile_parameters": tuned_parameters(),
                   "layers": len(layers), "gates": sum(len(l) for l in layers),
                   "max_weight": w.max_weight, "min_coeff": w.min_coeff, "damping": w.damping,
                   "peak_terms": peak, "term_gate_updates_per_step": updates // max(steps, 1),
                   "expectation": evs[-1] if evs else None},
        "roofline": roofline_of(prof, w, len(layers)),
        "phases_ms": {k: round(v["ms"], 3) for k, v in prof.items() if k != "sparse_gate"},
        "phases_share": {k: round(v["ms"] / total_phase_ms, 4) if total_phase_ms else 0.0 for k, v in prof.items()
                         if k != "sparse_gate"},
        "host_ms": {"enqueue": round(host["enqueue_ms"], 3), "wait_for_counts": round(host["wait_ms"], 3),
                    "gates": host["gates"]},
        "paths_per_step": {k: ctr[k] // max(steps, 1) for k in
                           ("small_gates", "sparse_gates", "sparse_pairs", "chained_gates", "sparse_fallbacks",
                            "sparse_syncs", "sparse_grows", "deferred")},
        "cpu_baseline": cpu,
    }
    eng.close()
    return out


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
    ap.add_argument("--no-majorana", action="store_true",
                    help="skip the Majorana UCJ leg (32 sites, 64 modes, mid-size stores)")
    ap.add_argument("--majorana-steps", type=int, default=2)
    ap.add_argument("--majorana-cpu-layers", type=int, default=96)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--sur-qubits", type=int, default=10)
    ap.add_argument("--sur-depth", type=int, default=4)
    ap.add_argument("--sur-sets", type=int, default=16384)
    ap.add_argument("--spec-instances", type=int, default=32,
                    help="batching as specified (depth U{4..16}): the bounded sample is drawn from this many")
    ap.add_argument("--spec-max-depth", type=int, default=9)
    ap.add_argument("--batch-depth-lo", type=int, default=1)
    ap.add_argument("--batch-depth-hi", type=int, default=5)
    ap.add_argument("--batch-grad-depth-hi", type=int, default=4)
    ap.add_argument("--no-batch-grad", action="store_true")
    ap.add_argument("--adjoint-checkpoint-every", type=int, default=0,
                    help="layers per checkpointed segment of the gradient legs (0: segments cut by the tape budget "
                         "only, so a tape that fits is recorded once and the sweep replays nothing)")
    ap.add_argument("--adjoint-tape-gib", type=float, default=64.0, help="the gradient legs' tape budget, GiB")
    ap.add_argument("--adjoint-check", type=int, default=32,
                    help="check this many seeded gradients against central differences (h=1e-4)")
    ap.add_argument("--shift-angles", type=int, default=4,
                    help="of those, check this many against parameter shift on the device")
    ap.add_argument("--shift-oracle-steps", type=int, default=4,
                    help="parameter shift with the oracle on the circuit's first k steps (0: off)")
    ap.add_argument("--shift-oracle-angles", type=int, default=8,
                    help="angles of that prefix checked against oracle parameter shift")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this engine runs on one GPU; the term store is not sharded across GPUs",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from propaq_amd import workloads

    w = workloads.config(args.config)
    layers = w.layers if args.layers is None else w.layers[:args.layers]
    head = measure(w, layers, args.steps, args.warmup, args.profile_every, args.path,
                   None if args.no_cpu else args.cpu_layers, args.verbose)
    mid = None
    if not args.no_majorana and args.config != 3:
        w3 = workloads.config(3)
        log(f"majorana leg: {w3.name}")
        mid = measure(w3, w3.layers, args.majorana_steps, 1, args.profile_every, args.path,
                      None if args.no_cpu else args.majorana_cpu_layers, args.verbose)
        mid = {"workload": w3.name, "value": mid["value"], "unit": "term-gate updates/s",
               "steps": args.majorana_steps, "warmup": 1, "ms_per_step": mid["ms_per_step"],
               "config": mid["config"], "roofline": mid["roofline"], "phases_ms": mid["phases_ms"],
               "host_ms": mid["host_ms"], "paths_per_step": mid["paths_per_step"],
               "cpu_baseline": mid["cpu_baseline"]}

    adj = None if args.no_adjoint else adjoint_leg(args)
    bat = None if args.no_batch else batch_leg(args)
    spec = None if args.no_batch else batch_spec_leg(args)
    sur = None if args.no_surrogate else surrogate_leg(args)

    out = {
        "metric": "term-gate updates/s (fwd) at fixed truncation; % HBM peak",
        "value": head["value"],
        "unit": "term-gate updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": head["config"],
        "roofline": head["roofline"],
        "phases_ms": head["phases_ms"],
        "phases_share": head["phases_share"],
        "phases_note": "sums over the sampled gates (every --profile-every-th gate) and all noise passes of the "
                       f"{args.steps} timed steps, not per step",
        "host_ms": head["host_ms"],
        "paths_per_step": head["paths_per_step"],
        "cpu_baseline": head["cpu_baseline"],
        "majorana": mid,
        "adjoint": adj,
        "batch": bat,
        "batch_as_specified": spec,
        "surrogate": sur,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
