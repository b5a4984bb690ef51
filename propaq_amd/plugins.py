"""Loader of the unchanged native plugin C ABI.

Noise (``crates/core/src/policy/native_noise.rs:12-28``)::

    uint32_t propaq_noise_abi_version(void);            // must be 1
    uint32_t propaq_noise_depends(void);                // optional: KEY=1, LAYER=2
    void*    propaq_noise_create(const char* json);     // optional, with destroy
    void     propaq_noise_destroy(void*);
    double   propaq_noise_factor(ctx, basis_kind, words, n_words, n_units, weight, layer, n_layers);
    int32_t  propaq_noise_factor_batch(..., const uint32_t* weights, ..., double* out, size_t n);

Truncator (``crates/core/src/policy/native_truncator.rs:13-33``): the same
pattern with ``propaq_truncator_keep`` (nonzero = keep) and the optional
``propaq_truncator_keep_batch`` (0 = success).

The engine keeps the terms on the device.  A weight-only model is tabulated
once (a layer-aware one once per layer) and the table is uploaded; a
key-aware model is evaluated here, on the host, over the keys exported from
the device, and the factors / keep flags go back to the device
(``crates/core/src/engine/noise_resolver.rs:78-131`` picks the same three
strategies).  ``words`` is NULL unless the plugin declared KEY, as in the
reference (:107-116).
"""

from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from .truncation import Truncator

DEPENDS_KEY, DEPENDS_LAYER, DEPENDS_KNOWN = 1, 2, 3
ABI_VERSION = 1
BATCH = 1024  # term_kernel.rs:93 KERNEL_BATCH

_U64P = ctypes.POINTER(ctypes.c_uint64)
_U32P = ctypes.POINTER(ctypes.c_uint32)
_DP = ctypes.POINTER(ctypes.c_double)


class _Plugin:
    prefix = ""
    kind = ""

    def __init__(self, path: str, config: Optional[str] = None):
        try:
            lib = ctypes.CDLL(path)
        except OSError as e:
            raise OSError(f"failed to load {self.kind} plugin '{path}': {e}") from e
        self._lib = lib
        p = self.prefix
        ver = self._sym(path, f"{p}_abi_version", ctypes.c_uint32, [])
        v = ver()
        if v != ABI_VERSION:
            raise ValueError(f"{self.kind} plugin '{path}' targets ABI version {v}, expected {ABI_VERSION}")
        dep = getattr(lib, f"{p}_depends", None)
        bits = 0
        if dep is not None:
            dep.restype, dep.argtypes = ctypes.c_uint32, []
            bits = int(dep())
            if bits & ~DEPENDS_KNOWN:
                raise ValueError(f"{self.kind} plugin '{path}' declares unknown depends bits {bits:#x}")
        self._depends = bits
        create = getattr(lib, f"{p}_create", None)
        destroy = getattr(lib, f"{p}_destroy", None)
        if (create is None) != (destroy is None):
            raise ValueError(f"{self.kind} plugin '{path}' must export both {p}_create and {p}_destroy, or neither")
        self._ctx = None
        self._destroy = None
        if create is not None:
            create.restype, create.argtypes = ctypes.c_void_p, [ctypes.c_char_p]
            destroy.restype, destroy.argtypes = None, [ctypes.c_void_p]
            self._ctx = create(config.encode() if config is not None else None)
            self._destroy = destroy
        self._bind(path)

    def _sym(self, path, name, res, args, required=True):
        f = getattr(self._lib, name, None)
        if f is None:
            if required:
                raise ValueError(f"{self.kind} plugin '{path}' does not export {name}")
            return None
        f.restype, f.argtypes = res, args
        return f

    def __del__(self):
        d = getattr(self, "_destroy", None)
        if d is not None:
            d(self._ctx)
            self._destroy = None

    @property
    def abi_version(self) -> int:
        return ABI_VERSION

    @property
    def depends(self) -> int:
        return self._depends

    @property
    def reads_key(self) -> bool:
        return bool(self._depends & DEPENDS_KEY)

    @property
    def reads_layer(self) -> bool:
        return bool(self._depends & DEPENDS_LAYER)

    def _words(self, words):
        if not self.reads_key or words is None:
            return None, 0
        a = np.ascontiguousarray(words, dtype=np.uint64)
        return a, a.size


class NativeNoiseModel(_Plugin):
    """Noise model backed by a native (C/Rust/AOT-Julia) plugin."""

    prefix, kind = "propaq_noise", "noise"

    def _bind(self, path):
        self._factor = self._sym(path, "propaq_noise_factor", ctypes.c_double,
                                 [ctypes.c_void_p, ctypes.c_uint32, _U64P, ctypes.c_size_t, ctypes.c_uint32,
                                  ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32])
        self._batch = self._sym(path, "propaq_noise_factor_batch", ctypes.c_int32,
                                [ctypes.c_void_p, ctypes.c_uint32, _U64P, ctypes.c_size_t, ctypes.c_uint32, _U32P,
                                 ctypes.c_uint32, ctypes.c_uint32, _DP, ctypes.c_size_t], required=False)

    def factor_term(self, basis_kind: int, words, n_units: int, weight: int, layer_index: int = 0,
                    n_layers: int = 0) -> float:
        a, nw = self._words(words)
        ptr = a.ctypes.data_as(_U64P) if a is not None else None
        return float(self._factor(self._ctx, basis_kind, ptr, nw, n_units, weight, layer_index, n_layers))

    def table(self, basis_kind: int, max_weight: int, layer_index: int = 0, n_layers: int = 0):
        """retabulate (noise_resolver.rs:56): factor for w in 0..=max_weight,
        words NULL, n_units = max_weight."""
        return [float(self._factor(self._ctx, basis_kind, None, 0, max_weight, w, layer_index, n_layers))
                for w in range(max_weight + 1)]

    def factors(self, basis_kind: int, keys: np.ndarray, weights: np.ndarray, n_units: int, layer_index: int,
                n_layers: int) -> np.ndarray:
        """Per-term factors in KERNEL_BATCH chunks (termsum.rs:340-371)."""
        n, W = keys.shape
        out = np.empty(n, dtype=np.float64)
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        weights = np.ascontiguousarray(weights, dtype=np.uint32)
        for s in range(0, n, BATCH):
            e = min(n, s + BATCH)
            kk = keys[s:e]
            ww = weights[s:e]
            oo = np.empty(e - s, dtype=np.float64)
            ok = False
            if self._batch is not None:
                wp = kk.ctypes.data_as(_U64P) if self.reads_key else None
                rc = self._batch(self._ctx, basis_kind, wp, W if self.reads_key else 0, n_units,
                                 ww.ctypes.data_as(_U32P), layer_index, n_layers, oo.ctypes.data_as(_DP), e - s)
                ok = rc == 0
            if not ok:
                for i in range(e - s):
                    wp = kk[i].ctypes.data_as(_U64P) if self.reads_key else None
                    oo[i] = self._factor(self._ctx, basis_kind, wp, W if self.reads_key else 0, n_units,
                                         int(ww[i]), layer_index, n_layers)
            out[s:e] = oo
        return out

    def __repr__(self):
        return f"NativeNoiseModel(<native plugin, depends={self._depends}>)"


class NativeTruncator(_Plugin, Truncator):
    """Truncation policy backed by a native (C/Rust/AOT-Julia) plugin."""

    prefix, kind = "propaq_truncator", "truncation"

    def _bind(self, path):
        self._keep = self._sym(path, "propaq_truncator_keep", ctypes.c_int32,
                               [ctypes.c_void_p, ctypes.c_uint32, _U64P, ctypes.c_size_t, ctypes.c_uint32,
                                ctypes.c_uint32, ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32])
        self._batch = self._sym(path, "propaq_truncator_keep_batch", ctypes.c_int32,
                                [ctypes.c_void_p, ctypes.c_uint32, _U64P, ctypes.c_size_t, ctypes.c_uint32, _U32P,
                                 _DP, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint8),
                                 ctypes.c_size_t], required=False)

    def keep_term(self, basis_kind: int, words, n_units: int, weight: int, coeff_magnitude: float,
                  layer_index: int = 0, n_layers: int = 0) -> bool:
        a, nw = self._words(words)
        ptr = a.ctypes.data_as(_U64P) if a is not None else None
        return bool(self._keep(self._ctx, basis_kind, ptr, nw, n_units, weight, float(coeff_magnitude),
                               layer_index, n_layers))

    def keep_flags(self, basis_kind: int, keys: np.ndarray, weights: np.ndarray, mags: np.ndarray, n_units: int,
                   layer_index: int, n_layers: int, batch: bool = True) -> np.ndarray:
        """Keep decisions; the batch entry point only in the reclaim sweep
        (native_truncator.rs:35-40), the scalar one per child."""
        n, W = keys.shape
        out = np.empty(n, dtype=np.uint8)
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        weights = np.ascontiguousarray(weights, dtype=np.uint32)
        mags = np.ascontiguousarray(mags, dtype=np.float64)
        for s in range(0, n, BATCH):
            e = min(n, s + BATCH)
            kk, ww, mm = keys[s:e], weights[s:e], mags[s:e]
            oo = np.empty(e - s, dtype=np.uint8)
            ok = False
            if batch and self._batch is not None:
                wp = kk.ctypes.data_as(_U64P) if self.reads_key else None
                rc = self._batch(self._ctx, basis_kind, wp, W if self.reads_key else 0, n_units,
                                 ww.ctypes.data_as(_U32P), mm.ctypes.data_as(_DP), layer_index, n_layers,
                                 oo.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), e - s)
                ok = rc == 0
            if not ok:
                for i in range(e - s):
                    wp = kk[i].ctypes.data_as(_U64P) if self.reads_key else None
                    oo[i] = 1 if self._keep(self._ctx, basis_kind, wp, W if self.reads_key else 0, n_units,
                                            int(ww[i]), float(mm[i]), layer_index, n_layers) else 0
            out[s:e] = oo
        return out

    def __repr__(self):
        return f"NativeTruncator(<native plugin, depends={self._depends}>)"
