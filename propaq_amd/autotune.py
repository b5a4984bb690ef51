"""Autotuner of the kernels' compile-time tile parameters.

The per-gate kernels take their tile shapes as compile-time constants
(``pq_kernels.h``):

* ``PQ_STORE_ITEMS``: outputs per thread of a store-merge tile at W = 4 (tile =
  256 x items; 7 uses 75 KiB of LDS, two tiles per CU);
* ``PQ_LOOKBACK_WIDTH``: predecessor statuses each lane reads per look-back step;
* ``PQ_MERGE_ITEMS``: outputs per thread of a class-bucket merge tile (a power
  of two: the generic child sort is a bitonic network over the same tile);
* ``PQ_BRANCH_ITEMS``: store rows per thread of a branch tile (the bucket
  histogram granularity);
* ``PQ_SPARSE_ITEMS``: store rows per thread of a sparse-gate gather block
  (block = 256 x items rows);
* ``PQ_JOIN_ITEMS``: list entries per thread of a sparse-gate join tile.

The store-merge, look-back, merge and branch parameters shape the sorted
chained kernels (recording, batched engines, keep kernels); the default
benchmark runs on the sparse gates, so the grid below varies the sparse
parameters and keeps the others at their recorded values.

``--build`` compiles one library per point of the grid (in parallel, on any
machine with hipcc); ``--measure`` runs the throughput benchmark's forward
workload once per build on the GPU (``PROPAQ_AMD_LIB`` selects the build),
rejects builds whose expectation value differs from the default build's in
any bit, and writes the fastest point to ``tuned.json`` beside this file.
``build_lib.build()`` compiles with the recorded point from then on.

    python -m propaq_amd.autotune --build      # CPU, hipcc
    python -m propaq_amd.autotune --measure    # on the MI355X
"""

from __future__ import annotations

import argparse
import concurrent.futures
import itertools
import json
import os
import subprocess
import sys
import time

from . import build_lib

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
TUNED = os.path.join(HERE, "tuned.json")
TUNE_DIR = os.path.join(HERE, "_lib", "tune")

GRID = {
    "PQ_STORE_ITEMS": [7],
    "PQ_LOOKBACK_WIDTH": [1],
    "PQ_MERGE_ITEMS": [4],
    "PQ_BRANCH_ITEMS": [4],
    "PQ_SPARSE_ITEMS": [2, 4, 8],
    "PQ_JOIN_ITEMS": [1, 2, 4],
}


def points():
    keys = list(GRID)
    for vals in itertools.product(*(GRID[k] for k in keys)):
        yield dict(zip(keys, vals))


def name_of(point: dict) -> str:
    return "_".join(f"{k[3:].lower()}{v}" for k, v in point.items())


def lib_of(point: dict) -> str:
    return os.path.join(TUNE_DIR, name_of(point), "libpropaq_amd.so")


def _build_one(point):
    t0 = time.perf_counter()
    build_lib.build(out_path=lib_of(point), defines=[f"{k}={v}" for k, v in point.items()])
    return name_of(point), time.perf_counter() - t0


def build_all(jobs: int) -> None:
    pts = list(points())
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        for name, dt in ex.map(_build_one, pts):
            print(f"built {name} ({dt:.0f} s)", flush=True)


def _bench(lib, steps: int, warmup: int, timeout: int, config: int = 2) -> dict:
    env = dict(os.environ)
    if lib:
        env["PROPAQ_AMD_LIB"] = lib
    cmd = [sys.executable, os.path.join(ROOT, "bench.py"), "--steps", str(steps), "--warmup", str(warmup),
           "--config", str(config), "--no-cpu", "--no-majorana", "--no-adjoint", "--no-batch", "--no-surrogate"]
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"benchmark failed ({r.returncode}): {r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def measure_all(steps: int, warmup: int, timeout: int, out_path: str = TUNED, config: int = 2) -> dict:
    ref = _bench(None, steps, warmup, timeout, config)
    ref_ev = ref["config"]["expectation"]
    print(f"default build: {ref['ms_per_step']:.1f} ms/step, <O> = {ref_ev!r}", flush=True)
    rows = []
    for point in points():
        lib = lib_of(point)
        if not os.path.exists(lib):
            raise FileNotFoundError(f"{lib}: run --build first")
        d = _bench(lib, steps, warmup, timeout, config)
        same = d["config"]["expectation"] == ref_ev
        row = {"defines": point, "ms_per_step": d["ms_per_step"], "value": d["value"],
               "roofline_kernel": d["roofline"]["kernel"], "roofline_kernel_us": d["roofline"]["avg_launch_ms"] * 1e3,
               "roofline_frac": d["roofline"]["frac"], "phases_ms": d.get("phases_ms"), "expectation_equal": same}
        rows.append(row)
        print(f"{name_of(point)}: {d['ms_per_step']:.1f} ms/step, {row['roofline_kernel']} "
              f"{row['roofline_kernel_us']:.1f} us, equal {same}", flush=True)
    ok = [r for r in rows if r["expectation_equal"]]
    if not ok:
        raise RuntimeError("no build reproduced the default build's expectation value")
    best = min(ok, key=lambda r: r["ms_per_step"])
    out = {"defines": best["defines"], "ms_per_step": best["ms_per_step"], "value": best["value"],
           "workload": ref["config"]["workload"], "default_ms_per_step": ref["ms_per_step"],
           "measured_at": time.strftime("%Y-%m-%d %H:%M:%S"), "candidates": rows}
    with open(out_path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"selected {name_of(best['defines'])}: {best['ms_per_step']:.1f} ms/step -> {out_path}", flush=True)
    return out


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--build", action="store_true", help="compile every grid point")
    ap.add_argument("--measure", action="store_true", help="time every grid point on the GPU, write tuned.json")
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--timeout", type=int, default=120, help="seconds per benchmark run")
    ap.add_argument("--out", default=TUNED, help="where --measure writes the selection (default: tuned.json)")
    ap.add_argument("--config", type=int, default=2, help="the benchmark workload to time (bench.py --config)")
    a = ap.parse_args()
    if not (a.build or a.measure):
        ap.error("give --build and/or --measure")
    if a.build:
        build_all(a.jobs)
    if a.measure:
        measure_all(a.steps, a.warmup, a.timeout, a.out, a.config)


if __name__ == "__main__":
    main()
