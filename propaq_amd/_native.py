"""ctypes binding of the C ABI in ``include/propaq_amd.h``.

This is the host-side stub a maintainer would add next to the reference's
PyO3 classes: every call goes straight to ``libpropaq_amd.so``.  There is no
fallback: if the library or a HIP device is missing, the calls raise.
"""

from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libpropaq_amd.so")

PQ_OK, PQ_ERR_ARG, PQ_ERR_DEVICE, PQ_ERR_MEMORY, PQ_ERR_STATE = 0, 1, 2, 3, 4
BASIS_PAULI, BASIS_MAJORANA = 0, 1

EXPORTED = [
    "pq_abi_version", "pq_last_error", "pq_device_count", "pq_engine_create", "pq_engine_destroy",
    "pq_engine_width", "pq_set_cutoff", "pq_set_defer_cliffords", "pq_seed", "pq_apply_rotation",
    "pq_apply_gates", "pq_scale_by_weight", "pq_scale_by_factors", "pq_reclaim", "pq_len", "pq_weights",
    "pq_expectation", "pq_expectation_noalloc", "pq_terms_below", "pq_export", "pq_export_raw", "pq_frame_is_identity",
    "pq_counters_get", "pq_counters_reset", "pq_key_bytes", "pq_synchronize", "pq_profile_enable",
    "pq_profile_get", "pq_set_differentiable", "pq_tape_begin", "pq_tape_end", "pq_tape_info", "pq_checkpoint_save",
    "pq_checkpoint_restore", "pq_checkpoint_free", "pq_adjoint_seed", "pq_adjoint_backward", "pq_adjoint_values",
    "pq_engine_create_batched", "pq_batch_info", "pq_seed_batched", "pq_apply_gates_batched",
    "pq_expectation_batched", "pq_instance_counts", "pq_retire_instances", "pq_readout_retire", "pq_adjoint_seed_batched",
    "pq_adjoint_backward_batched", "pq_freeze_instances", "pq_set_keep_kernel", "pq_set_layer",
    "pq_set_small_runs", "pq_set_sparse", "pq_set_sparse_budget", "pq_set_pairs", "pq_surrogate_create", "pq_surrogate_destroy", "pq_surrogate_width",
    "pq_surrogate_set_truncation", "pq_surrogate_seed", "pq_surrogate_apply", "pq_surrogate_len",
    "pq_surrogate_compile", "pq_model_create", "pq_model_destroy", "pq_model_info", "pq_model_export",
    "pq_model_evaluate", "pq_model_set_profiling", "pq_model_timing", "pq_surrogate_apply_logged",
    "pq_surrogate_counters", "pq_model_set_batch_limit", "pq_surrogate_set_compaction",
]

# pq_keep_batch_fn: the shape of the truncator plugin's propaq_truncator_keep_batch
KEEP_BATCH_FN = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64),
                                 ctypes.c_size_t, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
                                 ctypes.POINTER(ctypes.c_double), ctypes.c_uint32, ctypes.c_uint32,
                                 ctypes.POINTER(ctypes.c_uint8), ctypes.c_size_t)

PHASES = ["branch", "child_sort", "store_merge", "noise_reclaim", "scale_in_place", "sparse_gather", "sparse_gate",
          "sparse_join"]


class SurrogateGateStats(ctypes.Structure):
    _fields_ = [("terms_rotated", ctypes.c_uint64), ("terms", ctypes.c_uint64), ("monomials_lo", ctypes.c_uint64),
                ("monomials_hi", ctypes.c_uint64), ("rotation_ms", ctypes.c_double),
                ("truncation_ms", ctypes.c_double)]


class Counters(ctypes.Structure):
    _fields_ = [
        ("visited", ctypes.c_uint64), ("emitted", ctypes.c_uint64), ("declined", ctypes.c_uint64),
        ("hits", ctypes.c_uint64), ("created", ctypes.c_uint64), ("dropped_weight", ctypes.c_uint64),
        ("discarded_coeff_l1", ctypes.c_double), ("discarded_coeff_max", ctypes.c_double),
        ("gates", ctypes.c_uint64), ("deferred", ctypes.c_uint64), ("term_gate_updates", ctypes.c_uint64),
        ("small_gates", ctypes.c_uint64), ("chained_gates", ctypes.c_uint64), ("generic_sorts", ctypes.c_uint64),
        ("sparse_gates", ctypes.c_uint64), ("sparse_fallbacks", ctypes.c_uint64), ("sparse_pairs", ctypes.c_uint64),
        ("sparse_syncs", ctypes.c_uint64), ("sparse_grows", ctypes.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Profile(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_uint64 * 8), ("ms", ctypes.c_double * 8), ("bytes", ctypes.c_double * 8),
                ("host_gates", ctypes.c_uint64), ("host_enqueue_ms", ctypes.c_double), ("host_wait_ms", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {ph: {"launches": int(self.launches[i]), "ms": float(self.ms[i]), "bytes": float(self.bytes[i])}
                for i, ph in enumerate(PHASES)}


_lib = None


def load():
    """Load the native library, raising if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    # PROPAQ_AMD_LIB selects another build of the same library (for example
    # one compiled with other tile parameters by build_lib.py --out ... -D ...)
    path = os.environ.get("PROPAQ_AMD_LIB", LIB_PATH)
    if not os.path.exists(path):
        raise ImportError(
            f"propaq_amd native library not built ({path}); run __graft_entry__.build() "
            "or python propaq_amd/build_lib.py")
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    U64, U64P = ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)
    DP = ctypes.POINTER(ctypes.c_double)
    I = ctypes.c_int
    sig = {
        "pq_abi_version": (ctypes.c_uint32, []),
        "pq_last_error": (ctypes.c_char_p, []),
        "pq_device_count": (I, [ctypes.POINTER(I)]),
        "pq_engine_create": (I, [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(P)]),
        "pq_engine_destroy": (None, [P]),
        "pq_engine_width": (I, [P, ctypes.POINTER(ctypes.c_uint32)]),
        "pq_set_cutoff": (I, [P, I, ctypes.c_uint32, I, ctypes.c_double, I, U64]),
        "pq_set_defer_cliffords": (I, [P, I]),
        "pq_seed": (I, [P, U64P, DP, U64]),
        "pq_apply_rotation": (I, [P, U64P, ctypes.c_double, U64P]),
        "pq_apply_gates": (I, [P, U64P, DP, U64, U64P]),
        "pq_scale_by_weight": (I, [P, DP, U64, I]),
        "pq_scale_by_factors": (I, [P, DP, U64]),
        "pq_reclaim": (I, [P, ctypes.POINTER(ctypes.c_uint8), U64]),
        "pq_len": (I, [P, U64P]),
        "pq_weights": (I, [P, ctypes.POINTER(ctypes.c_uint32), U64]),
        "pq_expectation": (I, [P, U64P, U64, DP]),
        "pq_expectation_noalloc": (I, [P, U64P, U64, DP]),
        "pq_terms_below": (I, [P, ctypes.c_double, U64P]),
        "pq_export": (I, [P, U64P, DP, U64, U64P]),
        "pq_export_raw": (I, [P, U64P, DP, U64, U64P]),
        "pq_frame_is_identity": (I, [P, ctypes.POINTER(I)]),
        "pq_counters_get": (I, [P, ctypes.POINTER(Counters)]),
        "pq_counters_reset": (I, [P]),
        "pq_key_bytes": (I, [P, U64P]),
        "pq_synchronize": (I, [P]),
        "pq_profile_enable": (I, [P, I]),
        "pq_profile_get": (I, [P, ctypes.POINTER(Profile)]),
        "pq_set_differentiable": (I, [P, I]),
        "pq_tape_begin": (I, [P]),
        "pq_tape_end": (I, [P]),
        "pq_tape_info": (I, [P, U64P, U64P]),
        "pq_checkpoint_save": (I, [P, ctypes.POINTER(ctypes.c_uint32)]),
        "pq_checkpoint_restore": (I, [P, ctypes.c_uint32]),
        "pq_checkpoint_free": (I, [P, ctypes.c_uint32]),
        "pq_adjoint_seed": (I, [P, U64P, U64]),
        "pq_adjoint_backward": (I, [P, DP, U64]),
        "pq_adjoint_values": (I, [P, DP, U64, U64P]),
        "pq_adjoint_seed_batched": (I, [P, U64P, U64]),
        "pq_adjoint_backward_batched": (I, [P, DP, U64]),
        "pq_freeze_instances": (I, [P, ctypes.POINTER(ctypes.c_uint8)]),
        "pq_set_keep_kernel": (I, [P, KEEP_BATCH_FN, P]),
        "pq_set_layer": (I, [P, ctypes.c_uint32, ctypes.c_uint32]),
        "pq_set_small_runs": (I, [P, I]),
        "pq_set_sparse": (I, [P, I]),
        "pq_set_sparse_budget": (I, [P, U64]),
        "pq_set_pairs": (I, [P, I]),
        "pq_engine_create_batched": (I, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(P)]),
        "pq_batch_info": (I, [P, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]),
        "pq_seed_batched": (I, [P, ctypes.POINTER(ctypes.c_uint32), U64P, DP, U64]),
        "pq_apply_gates_batched": (I, [P, U64P, DP, U64, U64P]),
        "pq_expectation_batched": (I, [P, U64P, U64, DP]),
        "pq_instance_counts": (I, [P, U64P]),
        "pq_retire_instances": (I, [P, ctypes.POINTER(ctypes.c_uint8)]),
        "pq_readout_retire": (I, [P, U64P, U64, ctypes.POINTER(ctypes.c_uint8), DP, U64P]),
        "pq_surrogate_create": (I, [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(P)]),
        "pq_surrogate_destroy": (None, [P]),
        "pq_surrogate_width": (I, [P, ctypes.POINTER(ctypes.c_uint32)]),
        "pq_surrogate_set_truncation": (I, [P, I, ctypes.c_uint32, I, ctypes.c_double, I, ctypes.c_uint32, I, U64,
                                            I]),
        "pq_surrogate_seed": (I, [P, U64P, DP, U64]),
        "pq_surrogate_apply": (I, [P, U64P, ctypes.POINTER(ctypes.c_int64), DP, U64, U64P, U64P]),
        "pq_surrogate_len": (I, [P, U64P]),
        "pq_surrogate_apply_logged": (I, [P, U64P, ctypes.POINTER(ctypes.c_int64), DP, U64,
                                          ctypes.POINTER(SurrogateGateStats)]),
        "pq_surrogate_counters": (I, [P, U64P, U64P, U64P, DP, DP, U64P]),
        "pq_surrogate_set_compaction": (I, [P, U64]),
        "pq_surrogate_compile": (I, [P, U64P, U64, ctypes.c_uint32, ctypes.POINTER(P)]),
        "pq_model_create": (I, [ctypes.c_uint32, U64, ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(ctypes.c_uint32),
                                U64P, U64P, DP, U64, DP, U64P, ctypes.POINTER(P)]),
        "pq_model_destroy": (None, [P]),
        "pq_model_info": (I, [P, U64P, U64P, U64P, U64P, U64P, ctypes.POINTER(ctypes.c_uint32)]),
        "pq_model_export": (I, [P, ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(ctypes.c_uint32), U64P, U64P, DP,
                                DP, U64P]),
        "pq_model_evaluate": (I, [P, DP, U64, U64, DP]),
        "pq_model_set_profiling": (I, [P, I]),
        "pq_model_set_batch_limit": (I, [P, U64]),
        "pq_model_timing": (I, [P, DP, DP, DP, U64P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(code: int):
    if code == PQ_OK:
        return
    msg = load().pq_last_error().decode(errors="replace")
    if code == PQ_ERR_ARG:
        raise ValueError(f"propaq_amd: {msg}")
    if code == PQ_ERR_MEMORY:
        raise MemoryError(f"propaq_amd: {msg}")
    raise RuntimeError(f"propaq_amd: {msg}")


def device_count() -> int:
    c = ctypes.c_int(0)
    _check(load().pq_device_count(ctypes.byref(c)))
    return c.value


def _u64p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def int_to_words(v: int, W: int) -> np.ndarray:
    return np.array([(v >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(W)], dtype=np.uint64)


def words_to_int(words) -> int:
    v = 0
    for i, w in enumerate(words):
        v |= int(w) << (64 * i)
    return v


class DeviceEngine:
    """One device-resident term sum (see ``pq_engine_create``), or
    ``n_instances`` of them sharing one store (``pq_engine_create_batched``)."""

    def __init__(self, basis_kind: int, n_units: int, n_instances: Optional[int] = None):
        L = load()
        h = ctypes.c_void_p()
        if n_instances is None:
            _check(L.pq_engine_create(int(basis_kind), int(n_units), ctypes.byref(h)))
        else:
            _check(L.pq_engine_create_batched(int(basis_kind), int(n_units), int(n_instances), ctypes.byref(h)))
        self._h = h
        self.kind, self.n_units = int(basis_kind), int(n_units)
        self.n_instances = n_instances
        w = ctypes.c_uint32(0)
        _check(L.pq_engine_width(self._h, ctypes.byref(w)))
        self.W = int(w.value)
        self.instance_shift = 64
        if n_instances is not None:
            ni, sh = ctypes.c_uint32(0), ctypes.c_uint32(0)
            _check(L.pq_batch_info(self._h, ctypes.byref(ni), ctypes.byref(sh)))
            self.instance_shift = int(sh.value)

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            load().pq_engine_destroy(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_cutoff(self, max_weight: Optional[int] = None, min_coeff: Optional[float] = None,
                   min_terms: Optional[int] = None):
        _check(load().pq_set_cutoff(self._h, int(max_weight is not None), int(max_weight or 0),
                                    int(min_coeff is not None), float(min_coeff or 0.0),
                                    int(min_terms is not None), int(min_terms or 0)))

    def set_defer_cliffords(self, on: bool):
        _check(load().pq_set_defer_cliffords(self._h, int(bool(on))))

    def set_compaction(self, min_nodes: int):
        _check(load().pq_surrogate_set_compaction(self._h, int(min_nodes)))

    def seed(self, keys: np.ndarray, coeffs: np.ndarray):
        k = np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1, self.W)
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        assert k.shape[0] == c.size
        _check(load().pq_seed(self._h, _u64p(k), _dp(c), c.size))

    def apply_gates(self, gens: np.ndarray, angles: np.ndarray) -> np.ndarray:
        g = np.ascontiguousarray(gens, dtype=np.uint64).reshape(-1, self.W)
        a = np.ascontiguousarray(angles, dtype=np.float64)
        assert g.shape[0] == a.size
        out = np.zeros(a.size, dtype=np.uint64)
        if a.size:
            _check(load().pq_apply_gates(self._h, _u64p(g), _dp(a), a.size, _u64p(out)))
        return out

    def apply_rotation(self, gen: np.ndarray, angle: float) -> int:
        g = np.ascontiguousarray(gen, dtype=np.uint64).reshape(self.W)
        n = ctypes.c_uint64(0)
        _check(load().pq_apply_rotation(self._h, _u64p(g), float(angle), ctypes.byref(n)))
        return int(n.value)

    def scale_by_weight(self, table, reclaim: bool = True):
        t = np.ascontiguousarray(table, dtype=np.float64)
        _check(load().pq_scale_by_weight(self._h, _dp(t), t.size, int(bool(reclaim))))

    def reclaim(self, keep: Optional[np.ndarray] = None):
        if keep is None:
            _check(load().pq_reclaim(self._h, None, 0))
        else:
            f = np.ascontiguousarray(keep, dtype=np.uint8)
            _check(load().pq_reclaim(self._h, f.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), f.size))

    def scale_by_factors(self, factors: np.ndarray):
        f = np.ascontiguousarray(factors, dtype=np.float64)
        _check(load().pq_scale_by_factors(self._h, _dp(f), f.size))

    def __len__(self) -> int:
        n = ctypes.c_uint64(0)
        _check(load().pq_len(self._h, ctypes.byref(n)))
        return int(n.value)

    def weights(self) -> np.ndarray:
        n = len(self)
        out = np.zeros(max(n, 1), dtype=np.uint32)
        _check(load().pq_weights(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), out.size))
        return out[:n]

    def expectation(self, diag_state: int = 0) -> float:
        nd = max(1, (self.n_units + 63) // 64)
        d = int_to_words(int(diag_state), nd)
        out = ctypes.c_double(0.0)
        _check(load().pq_expectation(self._h, _u64p(d), nd, ctypes.byref(out)))
        return float(out.value)

    def terms_below(self, thr: float) -> int:
        n = ctypes.c_uint64(0)
        _check(load().pq_terms_below(self._h, float(thr), ctypes.byref(n)))
        return int(n.value)

    def export(self, frame: bool = True):
        """(keys uint64[n, W], coeffs float64[n]) in store order."""
        n = len(self)
        keys = np.zeros((max(n, 1), self.W), dtype=np.uint64)
        coeffs = np.zeros(max(n, 1), dtype=np.float64)
        got = ctypes.c_uint64(0)
        fn = load().pq_export if frame else load().pq_export_raw
        _check(fn(self._h, _u64p(keys), _dp(coeffs), max(n, 1), ctypes.byref(got)))
        m = int(got.value)
        return keys[:m], coeffs[:m]

    def frame_is_identity(self) -> bool:
        v = ctypes.c_int(0)
        _check(load().pq_frame_is_identity(self._h, ctypes.byref(v)))
        return bool(v.value)

    def counters(self) -> dict:
        c = Counters()
        _check(load().pq_counters_get(self._h, ctypes.byref(c)))
        return c.as_dict()

    def reset_counters(self):
        _check(load().pq_counters_reset(self._h))

    def key_bytes(self) -> int:
        n = ctypes.c_uint64(0)
        _check(load().pq_key_bytes(self._h, ctypes.byref(n)))
        return int(n.value)

    def synchronize(self):
        _check(load().pq_synchronize(self._h))

    def profile_enable(self, every: int = 1):
        """Time phases with HIP events on every ``every``-th gate (0 or
        False: off; True or 1: every gate)."""
        _check(load().pq_profile_enable(self._h, int(every)))

    def profile(self) -> dict:
        p = Profile()
        _check(load().pq_profile_get(self._h, ctypes.byref(p)))
        return p.as_dict()

    def profile_host(self) -> dict:
        """Host wall time of the profiled gates: enqueueing and waiting for
        row counts."""
        p = Profile()
        _check(load().pq_profile_get(self._h, ctypes.byref(p)))
        return {"gates": int(p.host_gates), "enqueue_ms": float(p.host_enqueue_ms), "wait_ms": float(p.host_wait_ms)}

    # -- batching --------------------------------------------------------
    def seed_batched(self, instances: np.ndarray, keys: np.ndarray, coeffs: np.ndarray):
        i = np.ascontiguousarray(instances, dtype=np.uint32)
        k = np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1, self.W)
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        assert i.size == k.shape[0] == c.size
        _check(load().pq_seed_batched(self._h, i.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), _u64p(k), _dp(c),
                                      c.size))

    def apply_gates_batched(self, gens: np.ndarray, angles: np.ndarray) -> np.ndarray:
        """angles: float64[n_gates, n_instances]."""
        g = np.ascontiguousarray(gens, dtype=np.uint64).reshape(-1, self.W)
        a = np.ascontiguousarray(angles, dtype=np.float64).reshape(g.shape[0], self.n_instances)
        out = np.zeros(g.shape[0], dtype=np.uint64)
        if g.shape[0]:
            _check(load().pq_apply_gates_batched(self._h, _u64p(g), _dp(a), g.shape[0], _u64p(out)))
        return out

    def _diag_words(self, diag_states):
        nd = max(1, (self.n_units + 63) // 64)
        if np.isscalar(diag_states):
            diag_states = [int(diag_states)] * self.n_instances
        return np.concatenate([int_to_words(int(v), nd) for v in diag_states]).astype(np.uint64), nd

    def expectation_batched(self, diag_states) -> np.ndarray:
        d, nd = self._diag_words(diag_states)
        out = np.zeros(self.n_instances, dtype=np.float64)
        _check(load().pq_expectation_batched(self._h, _u64p(d), nd, _dp(out)))
        return out

    def adjoint_seed_batched(self, diag_states):
        """Seed every row against its instance's basis state."""
        d, nd = self._diag_words(diag_states)
        _check(load().pq_adjoint_seed_batched(self._h, _u64p(d), nd))

    def adjoint_backward_batched(self, n_gates: int) -> np.ndarray:
        """d<O_b>/dtheta of every recorded gate and instance: [n_gates, n_instances]."""
        g = np.zeros((max(int(n_gates), 1), self.n_instances), dtype=np.float64)
        _check(load().pq_adjoint_backward_batched(self._h, _dp(g), int(n_gates)))
        return g[:int(n_gates)]

    def instance_counts(self) -> np.ndarray:
        out = np.zeros(self.n_instances, dtype=np.uint64)
        _check(load().pq_instance_counts(self._h, _u64p(out)))
        return out

    def set_keep_kernel(self, keep):
        """keep(basis_kind, keys uint64[n, W], weights uint32[n], magnitudes
        float64[n], n_units, layer_index, n_layers) -> uint8[n] replaces the
        emit gate's coefficient test and the reclaim (None removes it)."""
        if keep is None:
            self._keep_cb = None
            _check(load().pq_set_keep_kernel(self._h, KEEP_BATCH_FN(), None))
            return

        def cb(ctx, kind, words, n_words, n_units, weights, mags, layer, n_layers, out, n):
            try:
                k = np.ctypeslib.as_array(words, shape=(n, n_words))
                w = np.ctypeslib.as_array(weights, shape=(n,))
                m = np.ctypeslib.as_array(mags, shape=(n,))
                res = np.asarray(keep(int(kind), k, w, m, int(n_units), int(layer), int(n_layers)), dtype=np.uint8)
                np.ctypeslib.as_array(out, shape=(n,))[:] = res
                return 0
            except Exception:  # noqa: BLE001 -- reported as a failed engine call
                return 1

        self._keep_cb = KEEP_BATCH_FN(cb)  # kept alive with the engine
        _check(load().pq_set_keep_kernel(self._h, self._keep_cb, None))

    def set_small_runs(self, on: bool):
        """Small stores apply runs of gates in one workgroup (default on)."""
        _check(load().pq_set_small_runs(self._h, int(bool(on))))

    def set_sparse(self, on: bool):
        """Larger stores apply gates in place, appending new rows (default on);
        off: every gate merges into a sorted copy (the chained kernels)."""
        _check(load().pq_set_sparse(self._h, int(bool(on))))

    def set_pairs(self, on: bool):
        """Two consecutive sparse gates with commuting generators in one pass
        (default on); off: one pass per gate."""
        _check(load().pq_set_pairs(self._h, int(bool(on))))

    def set_sparse_budget(self, nbytes: int):
        """At most nbytes of device memory for the sparse gates' lists and
        tables (0: the device's free memory); gates beyond it run chained."""
        _check(load().pq_set_sparse_budget(self._h, int(nbytes)))

    def set_layer(self, layer_index: int, n_layers: int):
        _check(load().pq_set_layer(self._h, int(layer_index), int(n_layers)))

    def freeze_instances(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.uint8)
        assert f.size == self.n_instances
        _check(load().pq_freeze_instances(self._h, f.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))

    def retire_instances(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.uint8)
        assert f.size == self.n_instances
        _check(load().pq_retire_instances(self._h, f.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))

    def readout_retire(self, diag_states, flags):
        """Expectation values and row counts of the flagged instances, then
        their rows removed: (values, counts), zero where not flagged."""
        d, nd = self._diag_words(diag_states)
        f = np.ascontiguousarray(flags, dtype=np.uint8)
        assert f.size == self.n_instances
        vals = np.zeros(self.n_instances, dtype=np.float64)
        counts = np.zeros(self.n_instances, dtype=np.uint64)
        _check(load().pq_readout_retire(self._h, _u64p(d), nd, f.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                        _dp(vals), _u64p(counts)))
        return vals, counts

    # -- differentiation -------------------------------------------------
    def set_differentiable(self, on: bool):
        _check(load().pq_set_differentiable(self._h, int(bool(on))))

    def tape_begin(self):
        _check(load().pq_tape_begin(self._h))

    def tape_end(self):
        _check(load().pq_tape_end(self._h))

    def tape_info(self):
        g, b = ctypes.c_uint64(0), ctypes.c_uint64(0)
        _check(load().pq_tape_info(self._h, ctypes.byref(g), ctypes.byref(b)))
        return int(g.value), int(b.value)

    def checkpoint_save(self) -> int:
        i = ctypes.c_uint32(0)
        _check(load().pq_checkpoint_save(self._h, ctypes.byref(i)))
        return int(i.value)

    def checkpoint_restore(self, cid: int):
        _check(load().pq_checkpoint_restore(self._h, int(cid)))

    def checkpoint_free(self, cid: int):
        _check(load().pq_checkpoint_free(self._h, int(cid)))

    def adjoint_seed(self, diag_state: int = 0):
        nd = max(1, (self.n_units + 63) // 64)
        d = int_to_words(int(diag_state), nd)
        _check(load().pq_adjoint_seed(self._h, _u64p(d), nd))

    def adjoint_backward(self, n_gates: int) -> np.ndarray:
        g = np.zeros(max(int(n_gates), 1), dtype=np.float64)
        _check(load().pq_adjoint_backward(self._h, _dp(g), int(n_gates)))
        return g[:int(n_gates)]

    def adjoint_values(self) -> np.ndarray:
        n = ctypes.c_uint64(0)
        _check(load().pq_adjoint_values(self._h, None, 0, ctypes.byref(n)))
        out = np.zeros(max(int(n.value), 1), dtype=np.float64)
        _check(load().pq_adjoint_values(self._h, _dp(out), out.size, ctypes.byref(n)))
        return out[:int(n.value)]


class SurrogateBuilder:
    """The symbolic build (``pq_surrogate_create``): seed, gates, compile."""

    def __init__(self, basis_kind: int, n_units: int):
        L = load()
        h = ctypes.c_void_p()
        _check(L.pq_surrogate_create(int(basis_kind), int(n_units), ctypes.byref(h)))
        self._h = h
        self.kind, self.n_units = int(basis_kind), int(n_units)
        w = ctypes.c_uint32(0)
        _check(L.pq_surrogate_width(self._h, ctypes.byref(w)))
        self.W = int(w.value)

    def close(self):
        if getattr(self, "_h", None):
            load().pq_surrogate_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_truncation(self, frequency=None, coefficient=None, weight=None, min_terms=None, simplify=False):
        _check(load().pq_surrogate_set_truncation(
            self._h, int(frequency is not None), int(min(frequency or 0, 0xFFFFFFFF)), int(coefficient is not None),
            float(coefficient or 0.0), int(weight is not None), int(weight or 0), int(min_terms is not None),
            int(min_terms or 0), int(bool(simplify))))

    def set_compaction(self, min_nodes: int):
        _check(load().pq_surrogate_set_compaction(self._h, int(min_nodes)))

    def seed(self, keys: np.ndarray, coeffs: np.ndarray):
        k = np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1, self.W)
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        assert k.shape[0] == c.size
        _check(load().pq_surrogate_seed(self._h, _u64p(k), _dp(c), c.size))

    def apply(self, gens: np.ndarray, param_index: np.ndarray, angles: np.ndarray, monomials: bool = False):
        """Live counts after each gate (and the monomial counts when asked)."""
        g = np.ascontiguousarray(gens, dtype=np.uint64).reshape(-1, self.W)
        p = np.ascontiguousarray(param_index, dtype=np.int64)
        a = np.ascontiguousarray(angles, dtype=np.float64)
        n = a.size
        assert g.shape[0] == n and p.size == n
        counts = np.zeros(max(n, 1), dtype=np.uint64)
        mono = np.zeros(2 * max(n, 1), dtype=np.uint64) if monomials else None
        _check(load().pq_surrogate_apply(self._h, _u64p(g), p.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                         _dp(a), n, _u64p(counts), _u64p(mono) if mono is not None else None))
        if mono is None:
            return counts[:n]
        return counts[:n], [int(mono[2 * i]) | (int(mono[2 * i + 1]) << 64) for i in range(n)]

    def apply_logged(self, gens: np.ndarray, param_index: np.ndarray, angles: np.ndarray) -> list:
        """Per-gate dicts: terms_rotated, terms, monomials, rotation_ms, truncation_ms."""
        g = np.ascontiguousarray(gens, dtype=np.uint64).reshape(-1, self.W)
        p = np.ascontiguousarray(param_index, dtype=np.int64)
        a = np.ascontiguousarray(angles, dtype=np.float64)
        n = a.size
        stats = (SurrogateGateStats * max(n, 1))()
        _check(load().pq_surrogate_apply_logged(self._h, _u64p(g), p.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                                _dp(a), n, stats))
        return [{"terms_rotated": int(x.terms_rotated), "terms": int(x.terms),
                 "monomials": int(x.monomials_lo) | (int(x.monomials_hi) << 64), "rotation_ms": x.rotation_ms,
                 "truncation_ms": x.truncation_ms} for x in stats[:n]]

    def counters(self) -> dict:
        v = [ctypes.c_uint64(0) for _ in range(4)]
        t = [ctypes.c_double(0) for _ in range(2)]
        _check(load().pq_surrogate_counters(self._h, *[ctypes.byref(x) for x in v[:3]],
                                            *[ctypes.byref(x) for x in t], ctypes.byref(v[3])))
        return {"visited": int(v[0].value), "emitted": int(v[1].value), "hits": int(v[2].value),
                "rotation_s": t[0].value, "truncation_s": t[1].value, "nodes": int(v[3].value)}

    def __len__(self) -> int:
        n = ctypes.c_uint64(0)
        _check(load().pq_surrogate_len(self._h, ctypes.byref(n)))
        return int(n.value)

    def compile(self, diag_state: int, n_params: int) -> "DeviceModel":
        nd = max(1, (self.n_units + 63) // 64)
        d = int_to_words(int(diag_state), nd)
        h = ctypes.c_void_p()
        _check(load().pq_surrogate_compile(self._h, _u64p(d), nd, int(n_params), ctypes.byref(h)))
        return DeviceModel(h)


class DeviceModel:
    """A compiled surrogate in HBM (``pq_model``)."""

    def __init__(self, handle):
        self._h = handle
        self._info()

    @classmethod
    def from_tape(cls, n_params: int, kinds, params, a, b, vals, overlaps, roots) -> "DeviceModel":
        kinds = np.ascontiguousarray(kinds, dtype=np.uint8)
        params = np.ascontiguousarray(params, dtype=np.uint32)
        a = np.ascontiguousarray(a, dtype=np.uint64)
        b = np.ascontiguousarray(b, dtype=np.uint64)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        overlaps = np.ascontiguousarray(overlaps, dtype=np.float64)
        roots = np.ascontiguousarray(roots, dtype=np.uint64)
        h = ctypes.c_void_p()
        _check(load().pq_model_create(
            int(n_params), kinds.size, kinds.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
            params.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), _u64p(a), _u64p(b), _dp(vals), overlaps.size,
            _dp(overlaps), _u64p(roots), ctypes.byref(h)))
        return cls(h)

    def _info(self):
        v = [ctypes.c_uint64(0) for _ in range(5)]
        npar = ctypes.c_uint32(0)
        _check(load().pq_model_info(self._h, *[ctypes.byref(x) for x in v], ctypes.byref(npar)))
        self.n_terms, self.n_ops, self.n_levels = int(v[0].value), int(v[1].value), int(v[2].value)
        self.n_monomials = int(v[3].value) | (int(v[4].value) << 64)
        self.n_params = int(npar.value)

    def close(self):
        if getattr(self, "_h", None):
            load().pq_model_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def export(self):
        n, t = self.n_ops, self.n_terms
        kinds = np.zeros(max(n, 1), dtype=np.uint8)
        params = np.zeros(max(n, 1), dtype=np.uint32)
        a = np.zeros(max(n, 1), dtype=np.uint64)
        b = np.zeros(max(n, 1), dtype=np.uint64)
        vals = np.zeros(max(n, 1), dtype=np.float64)
        ov = np.zeros(max(t, 1), dtype=np.float64)
        roots = np.zeros(max(t, 1), dtype=np.uint64)
        _check(load().pq_model_export(self._h, kinds.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                      params.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), _u64p(a), _u64p(b),
                                      _dp(vals), _dp(ov), _u64p(roots)))
        return kinds[:n], params[:n], a[:n], b[:n], vals[:n], ov[:t], roots[:t]

    def evaluate(self, params: np.ndarray) -> np.ndarray:
        """params float64[n_sets, stride >= n_params] -> float64[n_sets]."""
        p = np.ascontiguousarray(params, dtype=np.float64)
        if p.ndim == 1:
            p = p.reshape(1, -1)
        n_sets, stride = p.shape
        out = np.zeros(max(n_sets, 1), dtype=np.float64)
        _check(load().pq_model_evaluate(self._h, _dp(p) if p.size else None, n_sets, stride, _dp(out)))
        return out[:n_sets]

    def set_batch_limit(self, max_sets: int):
        _check(load().pq_model_set_batch_limit(self._h, int(max_sets)))

    def set_profiling(self, on: bool):
        _check(load().pq_model_set_profiling(self._h, int(bool(on))))

    def timing(self) -> dict:
        t, s, by = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_double(0)
        n = ctypes.c_uint64(0)
        _check(load().pq_model_timing(self._h, ctypes.byref(t), ctypes.byref(s), ctypes.byref(by), ctypes.byref(n)))
        return {"tape_ms": t.value, "sum_ms": s.value, "tape_bytes": by.value, "launches": int(n.value)}
