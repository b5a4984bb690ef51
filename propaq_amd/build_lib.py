"""Build recipe of the native library ``propaq_amd/_lib/libpropaq_amd.so``.

The HIP kernels and the C++ host engine are compiled for gfx950 with hipcc.
``-ffp-contract=off`` keeps every ``a*b + c`` a separate multiply and add, as
the reference's Rust FP64 arithmetic rounds, so coefficients match bit for bit.
"""

import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "libpropaq_amd.so")
SOURCES = ["pq_kernels.hip", "pq_engine.cpp", "pq_tape.hip", "pq_surrogate.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r


TUNED = os.path.join(HERE, "tuned.json")


def tuned_defines():
    """The tile parameters propaq_amd/autotune.py selected on the GPU (none
    until it has run: the defaults of pq_kernels.h apply)."""
    if not os.path.exists(TUNED):
        return []
    with open(TUNED) as fh:
        return [f"{k}={v}" for k, v in json.load(fh)["defines"].items()]


def build(verbose: bool = False, out_path: str = LIB_PATH, defines=None) -> str:
    """Compile the library to out_path; defines (e.g. "PQ_STORE_ITEMS=2")
    override the compile-time tile parameters of pq_kernels.h.  Without
    defines the tuned ones are used."""
    if defines is None:
        defines = tuned_defines()
    out_dir = os.path.dirname(os.path.abspath(out_path))
    os.makedirs(out_dir, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(out_dir, src + ".o")
        cmd = [HIPCC, f"--offload-arch={ARCH}", *COMMON, *[f"-D{d}" for d in defines], "-c",
               os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd.insert(1, "-xhip")
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    tmp = out_path + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out_path)
    for o in objs:
        os.remove(o)
    return out_path


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=LIB_PATH)
    ap.add_argument("-D", dest="defines", action="append", default=[])
    a = ap.parse_args()
    print(build(verbose=True, out_path=a.out, defines=a.defines or None))
